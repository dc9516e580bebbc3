"""GPU parity at the workloads' full integration lengths, and the dispatcher
paths that carry a spectrum across several raw buffers.

Spectra are compared with the FP64 oracle (normwise relative error).  The
product path accumulates with Kahan-compensated FP32 and must stay within
1e-5; the plain-FP32 accumulation (SWSPEC_FLAG_NO_KAHAN) and the FP32
restatement of the reference core (plain FP32 sums, as TaskCoreIPP.cpp:
438-447) are measured beside it and reported, with a loose bound of 1e-4
that only catches gross errors.  Phase-cal folded vectors: <= 1e-6 against
the FP64 fold with the reference's FP32 rotator; tones <= 1e-6.

Workloads (seeded, generated on the device):
  32000-point Hann, 1 channel 2-bit raw, 1600 FFTs (51.2M samples), tone on bin 4800
  3.2M-point Cosine2, 4-channel Mk5B payload, 100 FFTs per channel, 16-tone
    phase-cal (the same bytes bench.py times)
  2^20-point Hann cross-pol pair (rho 0.2 at 30 degrees), 200 FFTs, XX, YY, X.Y*
  the 3.2M-point payload dispatched as TaskDispatcher does with
    MaxSourceBufferMB=32: K = 4 FFTs per buffer, 25 buffers per spectrum,
    combineResults after each (TaskDispatcher.cpp:143-171), host and device
    combine; the written spectrum is (1/K) sum over all 100 FFTs (TaskCoreIPP.cpp:124)
  the C++ TaskCoreHIP adapter run by oracle/_ref/taskcorehip_drv against the
    Python dispatcher, output files compared byte for byte
"""
import os
import subprocess
import threading

import numpy as np
import pytest

from gpu_util import oracle_settings, oracle_split, relerr, say

pytestmark = pytest.mark.gpu
TOL = 1e-5
LOOSE = 1e-4
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(hip, s, d_in, channels=None, flags=0):
    """One run of a core on device-resident inputs -> host arrays."""
    import torch
    core = hip.TaskCoreHIP(0)
    core.prepare(s, channels=channels, flags=flags)
    nj = core.njobs
    ssb = s.fft_ssb_points
    T = s.pcal_tonebins if s.extract_pcal else 0
    d_a = [torch.full((ssb,), float("nan"), device="cuda") for _ in range(nj)]
    d_x = [torch.full((2 * ssb,), float("nan"), device="cuda")] if s.num_xpols else None
    d_p = [torch.full((2 * T,), float("nan"), device="cuda") for _ in range(nj)] if T else None
    core.run_device(d_in, [int(t.numel()) for t in d_in], d_a, d_x, d_p)
    assert core.join() == 1
    auto = [a.cpu().numpy().astype(np.float64) for a in d_a]
    xp = []
    if d_x:
        v = d_x[0].cpu().numpy().astype(np.float64)
        xp = [v[0::2] + 1j * v[1::2]]
    pc = []
    if d_p:
        for p in d_p:
            v = p.cpu().numpy().astype(np.float64)
            pc.append(v[0::2] + 1j * v[1::2])
    tones = [core.pcal_tones(j, 0) for j in range(nj)] if T else []
    core.finalize()
    return auto, xp, pc, tones


def _bg(fn):
    """Run fn in a thread; returns a getter for its result."""
    box = {}

    def work():
        box["r"] = fn()

    t = threading.Thread(target=work)
    t.start()

    def get():
        t.join()
        return box["r"]
    return get


# ------------------------------------------------------------- 32000 x 1600

def test_full_length_32000pt_1600ffts(hip, oracle):
    import torch
    from swspec_amd import workloads as W
    from swspec_amd._lib import FLAG_NO_KAHAN
    n, nf, fs = 32000, 1600, 32e6
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=1.6, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=16e6, BitsPerSample=2, SourceChannels=1)
    assert s.core_overlapped_ffts == nf and s.averaged_ffts == nf
    amp = 4.0 * np.sqrt(3.75 / n)          # 10 dB per-FFT bin SNR under a Hann window
    say("full length 32000 x 1600: generating")
    raw = W.make_ring(torch, 1601, "raw1", n * nf, fs, [[(amp, 4.8e6, 0.0)]])
    host = raw.cpu().numpy()

    def mk(k):
        return oracle.derive(fft_points=n, fft_integ_seconds=n * (k + 0.25) / fs, fft_overlap_factor=1,
                             window="Hann", bandwidth_hz=16e6, bits_per_sample=2, source_channels=1,
                             max_source_buffer_mb=1 << 16)
    f32 = _bg(lambda: oracle.domaths(oracle_settings(oracle, s), [host], 32))
    a64, _, _ = oracle_split(oracle, mk, [host], nf, n // 4)
    kah, _, _, _ = _run(hip, s, [raw])
    plain, _, _, _ = _run(hip, s, [raw], flags=FLAG_NO_KAHAN)
    m32, r32, _, _ = f32()
    assert m32 == 1
    ek, ep, er = relerr(kah[0], a64[0]), relerr(plain[0], a64[0]), relerr(r32[0][0], a64[0])
    say("  vs FP64: Kahan %.3e  plain FP32 %.3e  FP32 restatement of the reference core %.3e" % (ek, ep, er))
    assert int(np.argmax(kah[0][1:-1])) + 1 == 4800
    assert ek <= TOL and ep <= LOOSE and er <= LOOSE


# ------------------------------------------------------------- 3.2M x 100, 4-channel Mk5B, phase-cal

N2, FS2, NF2 = 3200000, 32e6, 100
_mk5b4 = {}


def _mk5b4_data(hip, oracle):
    """The 4-channel payload (bench.py's seed), its FP64 oracle spectra and
    phase-cal folds per channel; computed once per module."""
    if _mk5b4:
        return _mk5b4
    import torch
    from swspec_amd import workloads as W
    say("3.2M x 100 x 4ch: generating")
    raw, _ = W.mk5b_payload(torch, 2024, N2 * NF2)
    host = raw.cpu().numpy()
    spectra, folds = [], []
    for ch in range(4):
        def mk(k, ch=ch):
            return oracle.derive(fft_points=N2, fft_integ_seconds=N2 * (k + 0.25) / FS2, fft_overlap_factor=1,
                                 window="Cosine2", bandwidth_hz=16e6, bits_per_sample=2, source_channels=4,
                                 source_format="Mk5B", use_channel_file1=ch + 1, extract_pcal=1,
                                 max_source_buffer_mb=1 << 16)
        a64, _, p64 = oracle_split(oracle, mk, [host], NF2, N2)
        spectra.append(a64[0])
        folds.append(p64[0])
    # FP64 fold with the reference's FP32 rotator (what the kernels apply), per FFT
    from concurrent.futures import ThreadPoolExecutor

    def fold(ch):
        acc = np.zeros(64)
        osc = oracle.derive(fft_points=N2, fft_integ_seconds=10.0, fft_overlap_factor=1, window="Cosine2",
                            bandwidth_hz=16e6, bits_per_sample=2, source_channels=4, source_format="Mk5B",
                            extract_pcal=1, max_source_buffer_mb=1024, use_channel_file1=ch + 1)
        for f in range(NF2):
            x = oracle.unpack(osc, host[f * N2:(f + 1) * N2], N2, ch).astype(np.float64)
            acc = oracle.pcal_fold64(x, FS2, 10e3, 32, 3200, rot32=True, acc=acc)
        return acc[0::2] + 1j * acc[1::2]
    with ThreadPoolExecutor(max_workers=4) as ex:
        fold32 = list(ex.map(fold, range(4)))
    _mk5b4.update(raw=raw, host=host, spectra=spectra, folds=folds, fold32=fold32)
    return _mk5b4


def _mk5b4_settings(hip, **over):
    k = dict(FFTpoints=N2, FFTIntegrationTimeSec=10.0, FFToverlapFactor=1, WindowType="Cosine2",
             BandwidthHz=16e6, BitsPerSample=2, SourceChannels=4, SourceFormat="Mk5B", ExtractPCal=True,
             PCalOffsetHz=10e3, MaxSourceBufferMB=1024)
    k.update(over)
    return hip.derive_settings(**k)


def test_full_length_3p2m_4ch_pcal(hip, oracle):
    from swspec_amd._lib import FLAG_NO_KAHAN
    d = _mk5b4_data(hip, oracle)
    s = _mk5b4_settings(hip)
    assert s.core_overlapped_ffts == NF2 and s.extract_pcal == 1 and s.pcal_tonebins == 32
    f32 = _bg(lambda: oracle.domaths(oracle_settings(oracle, s), [d["host"]], 32))
    kah, _, pc, tones = _run(hip, s, [d["raw"]], channels=[0, 1, 2, 3])
    plain, _, pcp, _ = _run(hip, s, [d["raw"]], channels=[0, 1, 2, 3], flags=FLAG_NO_KAHAN)
    for ch in range(4):
        ek, ep = relerr(kah[ch], d["spectra"][ch]), relerr(plain[ch], d["spectra"][ch])
        epc = relerr(pc[ch], d["fold32"][ch])
        epc64 = relerr(pc[ch], d["folds"][ch])
        acc = np.stack([d["fold32"][ch].real, d["fold32"][ch].imag], 1).reshape(-1)
        tr = oracle.pcal_tones64(acc, len(tones[ch]), 1e6, FS2)
        et = relerr(tones[ch], tr)
        say("  ch%d vs FP64: Kahan %.3e plain FP32 %.3e | pcal %.3e (vs FP64 rotator %.3e) tones %.3e" % (
            ch, ek, ep, epc, epc64, et))
        assert ek <= TOL and ep <= LOOSE
        assert epc <= 1e-6 and epc64 <= 1e-5 and et <= 1e-6
        assert np.array_equal(pc[ch], pcp[ch])   # the phase-cal fold does not depend on the accumulator mode
    m32, r32, _, _ = f32()
    er = relerr(r32[0][0], d["spectra"][0])
    say("  ch0 FP32 restatement of the reference core vs FP64: %.3e" % er)
    assert m32 == 1 and er <= LOOSE


# ------------------------------------------------------------- 2^20 cross-pol pair x 200

def test_full_length_2p20_xpol_200ffts(hip, oracle):
    import torch
    from swspec_amd import workloads as W
    from swspec_amd._lib import FLAG_MFMA64, FLAG_NO_KAHAN
    n, nf, fs = 1 << 20, 200, 64e6
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * (nf + 0.25) / fs, FFToverlapFactor=1,
                            WindowType="Hann", BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1,
                            DoCrossPolarization=True, MaxSourceBufferMB=4096, num_sources=2)
    assert s.core_overlapped_ffts == nf and s.num_xpols == 1
    say("cross-pol 2^20 x 200: generating")
    ra, rb = W.xpol_pair(torch, 3, n, nf, rho=0.2, phase_deg=30.0)
    ha, hb = ra.cpu().numpy(), rb.cpu().numpy()

    def mk(k):
        return oracle.derive(fft_points=n, fft_integ_seconds=n * (k + 0.25) / fs, fft_overlap_factor=1,
                             window="Hann", bandwidth_hz=fs / 2, bits_per_sample=2, source_channels=1,
                             max_source_buffer_mb=1 << 16, num_sources=2, do_xpol=1)
    f32 = _bg(lambda: oracle.domaths(oracle_settings(oracle, s), [ha, hb], 32))
    a64, x64, _ = oracle_split(oracle, mk, [ha, hb], nf, n // 4)
    ref = [a64[0], a64[1], x64[0]]
    # X.Y* = rho e^{+i 30deg} |X|^2 in band (Y's positive frequencies carry e^{-i phi})
    band = slice(n // 40, n // 2 - n // 40)
    ang = np.degrees(np.angle(x64[0][band].sum()))
    say("  X.Y* phase over the band %.2f deg" % ang)
    assert abs(ang - 30.0) < 2.0
    for name, fl in (("Kahan FP32", 0), ("plain FP32", FLAG_NO_KAHAN), ("MFMA64 stage + Kahan", FLAG_MFMA64)):
        a, x, _, _ = _run(hip, s, [ra, rb], flags=fl)
        errs = [relerr(a[0], ref[0]), relerr(a[1], ref[1]), relerr(x[0], ref[2])]
        say("  %-22s vs FP64: XX %.3e YY %.3e X.Y* %.3e" % (name, *errs))
        assert max(errs) <= (LOOSE if fl == FLAG_NO_KAHAN else TOL)
    m32, r32, rx32, _ = f32()
    errs = [relerr(r32[0][0], ref[0]), relerr(r32[1][0], ref[1]), relerr(rx32[0][0], ref[2])]
    say("  FP32 restatement of the reference core vs FP64: XX %.3e YY %.3e X.Y* %.3e" % tuple(errs))
    assert m32 == 1 and max(errs) <= LOOSE


# ------------------------------------------------------------- dispatcher: spectrum over 25 buffers

def _mk5b4_file(tmp_path_factory, d):
    p = tmp_path_factory.getbasetemp() / "mk5b4_payload.raw"
    if not p.exists():
        d["host"].tofile(str(p))
    return str(p)


def test_dispatcher_combine_25_buffers_host_and_device(hip, oracle, tmp_path_factory):
    """MaxSourceBufferMB=32: K = 4, 25 buffers per spectrum.  The Python
    TaskDispatcher (host Buffers, run/join/combineResults per buffer, then
    write) and a device loop (run_device/join/combine_device per buffer)
    give the same spectrum and phase-cal bit for bit, and both equal the
    oracle's (1/K) x sum over all 100 FFTs and the sum of all folds."""
    import torch
    from swspec_amd import dispatcher, sinks
    d = _mk5b4_data(hip, oracle)
    s = _mk5b4_settings(hip, MaxSourceBufferMB=32, UseFile1Channel=1)
    assert (s.core_overlapped_ffts, s.max_buffers_per_spectrum, s.max_spectra_per_buffer) == (4, 25, 0)
    assert s.rawbuf_size == 4 * N2
    path = _mk5b4_file(tmp_path_factory, d)
    prefix = str(tmp_path_factory.getbasetemp() / "py_mk5b4")
    nsp, files = dispatcher.run_files(s, [path], prefix)
    assert nsp == 1
    spec = sinks.read_spectra(files[0], N2)
    assert spec.shape == (1, N2 // 2 + 1)
    pv = np.fromfile(files[1], dtype="<f4")
    assert pv.size == 64
    host_pc = pv[0::2].astype(np.float64) + 1j * pv[1::2]
    want = d["spectra"][0] * (NF2 / 4)          # (1/K) sum over all FFTs, K = 4
    e, ep = relerr(spec[0], want), relerr(host_pc, d["folds"][0])
    say("  host combine over 25 buffers: spectrum vs FP64 %.3e, phase-cal %.3e" % (e, ep))
    assert e <= TOL and ep <= TOL
    # device-resident loop with combine_device
    core = hip.TaskCoreHIP(0)
    core.prepare(s)
    ssb, T = s.fft_ssb_points, s.pcal_tonebins
    comb = torch.full((2 * ssb,), 5.0, device="cuda")    # peak size, as the dispatcher allocates
    combp = torch.full((2 * T,), 5.0, device="cuda")
    hip.TaskCoreHIP.reset_device(comb)
    hip.TaskCoreHIP.reset_device(combp)
    d_a = torch.empty(ssb, device="cuda")
    d_p = torch.empty(2 * T, device="cuda")
    for b in range(s.max_buffers_per_spectrum):
        chunk = d["raw"][b * s.rawbuf_size:(b + 1) * s.rawbuf_size]
        core.run_device([chunk], [s.rawbuf_size], [d_a], None, [d_p])
        assert core.join() == 1
        ol, pl = core.combine_device([comb], [combp])
        assert ol == [4 * ssb] and pl == [8 * T]
    core.finalize()
    dev = comb[:ssb].cpu().numpy()
    dpc = combp.cpu().numpy()
    assert np.array_equal(dev, spec[0].astype(np.float32))
    assert np.array_equal(dpc, pv)


def test_cpp_adapter_driver_matches_python_dispatcher(hip, oracle, tmp_path_factory):
    """The C++ TaskCoreHIP (swspec_amd/host/TaskCoreHIP.cpp) driven by
    oracle/_ref/taskcorehip_drv -- the reference's Buffer and FileSink, the
    dispatcher loop for one core -- writes the same bytes as the Python
    dispatcher: (a) the 25-buffer combine case above, (b) two 1-channel
    sources with cross-power and phase-cal written per buffer (two spectra
    per buffer, three buffers)."""
    from swspec_amd import dispatcher, synth
    from oracle import oracle as O
    drv = O.REF_ADAPTER_DRV
    assert os.path.exists(drv), "oracle/_ref/taskcorehip_drv is built by __graft_entry__.build()"
    d = _mk5b4_data(hip, oracle)
    base = tmp_path_factory.getbasetemp()

    def both(keys, inputs, name):
        s = hip.derive_settings(num_sources=len(inputs), **keys)
        npy, pfiles = dispatcher.run_files(s, inputs, str(base / ("py_" + name)))
        ini = []
        for k, v in keys.items():
            if k == "WindowType":
                v = {"None": 0, "Cosine": 1, "Cosine2": 2, "Hamming": 3, "Hann": 4, "Blackman": 5}[v]
            elif k == "SourceFormat":
                v = {"RawSigned": 0, "RawUnsigned": 1, "Mk5B": 2}[v]
            elif isinstance(v, bool):
                v = int(v)
            ini.append("%s=%s" % (k, v))
        cmd = [drv, str(base / ("cc_" + name)), inputs[0], inputs[1] if len(inputs) > 1 else "-"] + ini
        say("  " + " ".join(cmd[1:]))
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        assert "spectra %d" % npy in r.stdout, r.stdout
        for pf in pfiles:
            cf = pf.replace("py_" + name, "cc_" + name)
            a, b = open(pf, "rb").read(), open(cf, "rb").read()
            assert len(a) > 0 and a == b, (pf, len(a), len(b))
        return s, pfiles

    keys = dict(FFTpoints=N2, FFTIntegrationTimeSec=10.0, FFToverlapFactor=1, WindowType="Cosine2",
                BandwidthHz=16e6, BitsPerSample=2, SourceChannels=4, SourceFormat="Mk5B", ExtractPCal=True,
                PCalOffsetHz=10e3, MaxSourceBufferMB=32, UseFile1Channel=3)
    both(keys, [_mk5b4_file(tmp_path_factory, d)], "combine")
    # (b) direct writes: 64000 points (20 rotator periods), overlap 2, 8 spectra per buffer
    n, fs = 64000, 32e6
    keys = dict(FFTpoints=n, FFTIntegrationTimeSec=n * 3 / fs, FFToverlapFactor=2, WindowType="Hann",
                BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, DoCrossPolarization=True, ExtractPCal=True,
                PCalOffsetHz=10e3, MaxSourceBufferMB=1)
    s = hip.derive_settings(num_sources=2, **keys)
    assert s.max_spectra_per_buffer == 8 and s.max_buffers_per_spectrum == 0 and s.num_xpols == 1
    assert s.extract_pcal == 1 and s.core_overlapped_ffts == 5
    inputs = []
    for k in range(2):
        raw = synth.raw_1ch(70 + k, 3 * s.rawbuf_size * 4, tones=[(0.1, 3.3e6, k)], fs=fs)
        p = str(base / ("direct_%d.raw" % k))
        raw.tofile(p)
        inputs.append(p)
    s, files = both(keys, inputs, "direct")
    spec = np.fromfile(files[0], dtype="<f4").reshape(-1, s.fft_ssb_points)
    assert spec.shape[0] == 3 * s.max_spectra_per_buffer
    # the first buffer's spectra against the oracle
    raw0 = [np.fromfile(p, dtype=np.uint8)[: s.rawbuf_size] for p in inputs]
    m, a64, x64, _ = oracle.domaths(oracle_settings(oracle, s), raw0, 64)
    k = s.max_spectra_per_buffer
    assert m == k
    e = relerr(spec[:k], a64[0])
    xs = np.fromfile(files[2], dtype="<f4")
    xs = (xs[0::2] + 1j * xs[1::2]).reshape(-1, s.fft_ssb_points)
    ex = relerr(xs[:k], x64[0])
    say("  direct mode vs FP64: spectra %.3e cross %.3e" % (e, ex))
    assert e <= TOL and ex <= TOL
