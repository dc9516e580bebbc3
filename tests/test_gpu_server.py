"""GPU tests of the batched/served engine (swspec_hip_server_*).

Each server job = one channel integrated over `nffts` non-overlapped FFTs,
scaled 1/nffts: the oracle's doMaths with K = nffts on the same bytes.
Tolerance: normwise relative error <= 1e-5 vs the FP64 oracle.
"""
import numpy as np
import pytest

from gpu_util import relerr, say
from swspec_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _oracle_mean(oracle, raw, n, nffts, fmt, nch, ch, window, fs, pcal=False):
    s = oracle.derive(fft_points=n, fft_integ_seconds=n * nffts / fs, fft_overlap_factor=1, window=window,
                      bandwidth_hz=fs / 2, bits_per_sample=2, source_channels=nch, source_format=fmt,
                      use_channel_file1=ch + 1, max_source_buffer_mb=1 << 16, extract_pcal=1 if pcal else 0)
    assert s.core_overlapped_ffts == nffts
    m, auto, _, pc = oracle.domaths(s, [raw], 64)
    assert m == 1
    return auto[0][0], (pc[0][0] if pcal else None), s


def _oracle_ring_mean(oracle, ring, n, nffts, fmt, nch, ch, window, fs):
    """A job reading a ring of R FFTs cyclically (FFT k from slot k mod R),
    averaged over nffts FFTs: sum over slots of (uses of the slot / nffts) x
    the slot's spectrum, each slot's spectrum from the FP64 oracle's doMaths
    with K = 1 (the combine identity, TaskCoreIPP.cpp:124)."""
    per = n // 4 if nch == 1 else n * (nch // 4)
    R = len(ring) // per
    s = oracle.derive(fft_points=n, fft_integ_seconds=n * 1.25 / fs, fft_overlap_factor=1, window=window,
                      bandwidth_hz=fs / 2, bits_per_sample=2, source_channels=nch, source_format=fmt,
                      use_channel_file1=ch + 1, max_source_buffer_mb=1 << 16)
    assert s.core_overlapped_ffts == 1
    m, auto, _, _ = oracle.domaths(s, [np.ascontiguousarray(ring[: R * per])], 64)
    assert m == R
    uses = np.bincount(np.arange(nffts) % R, minlength=R) / nffts
    return (uses[:, None] * auto[0]).sum(axis=0)


def _mkjob(kind, n, seed, nffts_data, fs):
    if kind == "raw1":
        return synth.raw_1ch(seed, n * nffts_data, tones=[(0.2, 0.21 * fs, 0.1)], fs=fs), "RawSigned", 1
    rng = np.random.default_rng(seed)
    tones = [synth.pcal_comb(fs, fs / 2, amp=0.15, rng=rng) for _ in range(4)]
    return synth.raw_multi(seed, 4, n * nffts_data, tones_per_ch=tones, fs=fs), "Mk5B", 4


def test_server_jobs_vs_oracle(hip, oracle):
    import torch
    from swspec_amd.server import Server
    fs = 32e6
    specs = [("raw1", 4096, 7, "Hann", 0), ("mk5b4", 20000, 5, "Cosine2", 2), ("raw1", 65536, 3, "Hamming", 0),
             ("mk5b4", 32000, 9, "Blackman", 1), ("raw1", 4096, 200, "Cosine2", 0), ("mk5b4", 320000, 3, "Cosine", 3)]
    srv = Server()
    outs, refs = [], []
    for i, (kind, n, nffts, win, ch) in enumerate(specs):
        say("server job", kind, n, nffts, win, ch)
        raw, fmt, nch = _mkjob(kind, n, 100 + i, nffts, fs)
        d_raw = torch.from_numpy(raw).cuda()
        out = torch.full((n // 2 + 1,), float("nan"), device="cuda")
        jid = srv.submit(n, nffts, d_raw, len(raw), out, window=win, source_format=fmt, source_channels=nch,
                         channel=ch, samplingfreq=fs, keep=d_raw)
        assert jid is not None
        ref, _, _ = _oracle_mean(oracle, raw, n, nffts, fmt, nch, ch, win, fs)
        outs.append(out)
        refs.append(ref)
    srv.run_batch(use_graph=False)
    for (spec, o, r) in zip(specs, outs, refs):
        e = relerr(o.cpu().numpy(), r)
        say("  ", spec, "err %.3e" % e)
        assert e <= TOL
    srv.close()


def test_server_ring_wrap_and_pcal(hip, oracle):
    """A job longer than its input ring reads slot k mod ring; phase-cal tones."""
    import torch
    from swspec_amd.server import Server
    fs, n, ring_ffts, nffts = 32e6, 32000, 3, 8
    raw, fmt, nch = _mkjob("mk5b4", n, 11, ring_ffts, fs)
    tiled = np.concatenate([raw] * (nffts // ring_ffts + 1))[: n * nffts]
    ref, pref, s = _oracle_mean(oracle, tiled, n, nffts, fmt, nch, 1, "Cosine2", fs, pcal=True)
    srv = Server()
    d_raw = torch.from_numpy(raw).cuda()
    out = torch.empty(n // 2 + 1, device="cuda")
    pc = torch.empty(2 * s.pcal_tonebins, device="cuda")
    nt = oracle.pcal_ntones(fs / 2, 1e6, 10e3)
    tones = torch.empty(2 * nt, dtype=torch.float64, device="cuda")
    srv.submit(n, nffts, d_raw, len(raw), out, window="Cosine2", source_format=fmt, source_channels=nch, channel=1,
               samplingfreq=fs, extract_pcal=True, d_out_pcal=pc, d_tones=tones)
    srv.run_batch(use_graph=False)
    e = relerr(out.cpu().numpy(), ref)
    p = pc.cpu().numpy()
    ep = relerr(p[0::2] + 1j * p[1::2], pref)
    acc = np.stack([pref.real, pref.imag], 1).reshape(-1)
    tr = oracle.pcal_tones64(acc, nt, 1e6, fs)
    t = tones.cpu().numpy()
    et = relerr(t[0::2] + 1j * t[1::2], tr)
    say("ring/pcal: spectrum %.3e pcal %.3e tones %.3e" % (e, ep, et))
    assert e <= TOL and ep <= 1e-5 and et <= 1e-5
    srv.close()


def _batch(torch, Server, use_graph, replays=1, seed=3):
    from swspec_amd import workloads as W
    streams = W.batch_workload(seed=seed, nstreams=12)
    for s in streams:
        s["nffts"] = min(s["nffts"], 40)
        s["fft_points"] = min(s["fft_points"], 320000)
    srv = Server()
    pool = W.RingPool(torch, ring_ffts=4)
    outs, total = W.submit_batch(torch, srv, streams, pool)
    srv.run_batch(use_graph=use_graph, replays=replays)
    res = [o.cpu().numpy() for o in outs]
    srv.close()
    return res, total


def test_server_graph_replay_matches_steps(hip):
    import torch
    from swspec_amd.server import Server
    say("graph vs steps")
    a, _ = _batch(torch, Server, use_graph=False)
    b, _ = _batch(torch, Server, use_graph=True, replays=3)
    for x, y in zip(a, b):
        assert np.isfinite(x).all()
        assert np.array_equal(x, y)


def test_result_same_alone_and_in_mix(hip, oracle):
    """Continuous batching: a job's result is the same alone and mixed with
    others, and every other job of the mix matches the oracle too.  Groups
    launch in plan-creation order on alternating streams, so when the
    4096-point group finishes (after 2 steps) the 65536-point job's batch
    moves to the other stream mid-run and must wait for its previous batch
    (counted in the server's cross-stream waits)."""
    import torch
    from swspec_amd.server import Server
    fs, n, nffts = 32e6, 320000, 12
    raw, fmt, nch = _mkjob("mk5b4", n, 21, nffts, fs)
    d_raw = torch.from_numpy(raw).cuda()
    srv = Server()
    alone = torch.empty(n // 2 + 1, device="cuda")
    srv.submit(n, nffts, d_raw, len(raw), alone, source_format=fmt, source_channels=4, channel=0, samplingfreq=fs)
    srv.run_batch(use_graph=False)
    # in a mix: same job admitted mid-run while other jobs of several sizes are
    # active; the inputs are rings of <= 8 FFTs read cyclically
    others = []
    for i, (m, k) in enumerate([(4096, 1500), (320000, 300), (65536, 3000), (320000, 7)]):
        r, f, c = _mkjob("raw1", m, 300 + i, min(k, 8), fs)
        d = torch.from_numpy(r).cuda()
        o = torch.full((m // 2 + 1,), float("nan"), device="cuda")
        srv.submit(m, k, d, len(r), o, source_format=f, source_channels=c, samplingfreq=fs, keep=d)
        others.append((r, m, k, o))
    srv.step()
    srv.step()
    mixed = torch.empty(n // 2 + 1, device="cuda")
    srv.submit(n, nffts, d_raw, len(raw), mixed, source_format=fmt, source_channels=4, channel=0, samplingfreq=fs)
    while srv.step() > 0:
        pass
    srv.sync()
    st = srv.stats()
    say("server stats", st)
    assert st["cross_stream_waits"] >= 1 and st["completed"] == st["admitted"] == 6
    ref, _, _ = _oracle_mean(oracle, raw, n, nffts, fmt, nch, 0, "Cosine2", fs)
    a, m = alone.cpu().numpy(), mixed.cpu().numpy()
    say("alone vs mix %.3e, alone vs oracle %.3e, mix vs oracle %.3e" % (relerr(m, a), relerr(a, ref), relerr(m, ref)))
    assert relerr(m, a) <= TOL and relerr(a, ref) <= TOL and relerr(m, ref) <= TOL
    for (r, mm, k, o) in others:
        e = relerr(o.cpu().numpy(), _oracle_ring_mean(oracle, r, mm, k, "RawSigned", 1, 0, "Cosine2", fs))
        say("  other job %d points x %d FFTs vs oracle %.3e" % (mm, k, e))
        assert e <= TOL
    srv.close()


def test_batch_64_streams_every_stream_vs_oracle(hip, oracle):
    """The 64-stream mixed-size batch (bench.py's workload: sizes 4096 ...
    3.2M, 10-5000 FFTs, raw 1-channel and Mk5B 4-channel, rings of 8 FFTs)
    captured into one hipGraph and replayed twice: every stream's spectrum
    within 1e-5 of the FP64 oracle (ring-weighted doMaths)."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from swspec_amd import workloads as W
    from swspec_amd.server import Server
    fs = 32e6
    streams = W.batch_workload(seed=4, nstreams=64)
    kinds = {(s["fft_points"], s["fmt"]) for s in streams}
    for need in [(1 << 20, "raw1"), (1 << 20, "mk5b4"), (3200000, "raw1"), (3200000, "mk5b4")]:
        assert need in kinds
    pool = W.RingPool(torch, ring_ffts=8)
    srv = Server(max_active=1024)
    outs, total = W.submit_batch(torch, srv, streams, pool)
    srv.run_batch(use_graph=True, replays=2)
    got = [o.cpu().numpy() for o in outs]
    srv.close()
    say("batch of %d streams, %.1f G channel-samples per replay" % (len(streams), total / 1e9))

    def ref(i):
        s = streams[i]
        ring = pool.get(s["fmt"], s["fft_points"], fs, round(s["tone"]), None, s["seed"]).cpu().numpy()
        if s["fmt"] == "raw1":
            return _oracle_ring_mean(oracle, ring, s["fft_points"], s["nffts"], "RawSigned", 1, 0, "Cosine2", fs)
        return _oracle_ring_mean(oracle, ring, s["fft_points"], s["nffts"], "Mk5B", 4, s["channel"], "Cosine2", fs)
    with ThreadPoolExecutor(max_workers=8) as ex:
        refs = list(ex.map(ref, range(len(streams))))
    worst = 0.0
    for s, g, r in zip(streams, got, refs):
        e = relerr(g, r)
        worst = max(worst, e)
        assert e <= TOL, (s, e)
    say("  worst stream vs FP64 oracle %.3e" % worst)


def test_server_mfma64_jobs_in_mix(hip, oracle):
    """A job flagged SWSPEC_FLAG_MFMA64 gets its own plan (same length,
    radix-64 MFMA column stage) and runs beside the FP32-plan copy of itself
    and jobs of other sizes: both within 1e-5 of the oracle and within 1e-6
    of each other."""
    import torch
    from swspec_amd._lib import FLAG_MFMA64
    from swspec_amd.server import Server
    fs, n, nffts = 32e6, 320000, 6
    raw, fmt, nch = _mkjob("mk5b4", n, 23, nffts, fs)
    d_raw = torch.from_numpy(raw).cuda()
    srv = Server()
    outs = {}
    for name, fl in (("fp32", 0), ("mfma", FLAG_MFMA64)):
        outs[name] = torch.empty(n // 2 + 1, device="cuda")
        assert srv.submit(n, nffts, d_raw, len(raw), outs[name], source_format=fmt, source_channels=4, channel=1,
                          samplingfreq=fs, flags=fl) is not None
    for i, (m, k) in enumerate([(65536, 9), (1 << 20, 3)]):
        r, f, c = _mkjob("raw1", m, 400 + i, k, fs)
        d = torch.from_numpy(r).cuda()
        o = torch.empty(m // 2 + 1, device="cuda")
        srv.submit(m, k, d, len(r), o, source_format=f, source_channels=c, samplingfreq=fs, keep=d,
                   flags=FLAG_MFMA64 if i else 0)
    while srv.step() > 0:
        pass
    srv.sync()
    ref, _, _ = _oracle_mean(oracle, raw, n, nffts, fmt, nch, 1, "Cosine2", fs)
    a, b = outs["fp32"].cpu().numpy(), outs["mfma"].cpu().numpy()
    say("fp32 vs oracle %.3e, mfma vs oracle %.3e, mfma vs fp32 %.3e" % (relerr(a, ref), relerr(b, ref), relerr(b, a)))
    assert relerr(a, ref) <= TOL and relerr(b, ref) <= TOL and relerr(b, a) <= 1e-6
    srv.close()


def test_server_admission_cap(hip):
    import torch
    from swspec_amd.server import Server
    srv = Server(max_active=2)
    d = torch.zeros(4096, dtype=torch.uint8, device="cuda")
    outs = [torch.empty(2049, device="cuda") for _ in range(3)]
    assert srv.submit(4096, 4, d, 1024, outs[0]) is not None
    assert srv.submit(4096, 4, d, 1024, outs[1]) is not None
    assert srv.submit(4096, 4, d, 1024, outs[2]) is None       # full: caller retries
    while srv.step() > 0:
        pass
    srv.sync()
    done = srv.poll()
    assert len(done) == 2
    assert srv.submit(4096, 4, d, 1024, outs[2]) is not None    # capacity freed
    srv.run_batch(use_graph=False)
    srv.close()
