"""GPU parity tests: TaskCoreHIP (HIP kernels via the C ABI) against the oracle.

Tolerances (floating point):
  spectra / cross-spectra : normwise relative error <= 1e-5 vs the FP64 oracle
                            and vs the FP32 restatement of the reference core
  decoded samples         : bit-exact
  phase-cal folded vector : relative error <= 1e-9 vs the FP64 fold with the
                            reference's FP32 rotator (FP64 accumulation on both)
"""
import numpy as np
import pytest

from gpu_util import oracle_settings, relerr, run_device, say
from swspec_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


# ---------------------------------------------------------------- decoders

@pytest.mark.parametrize("fmt,nch,incr", [("RawSigned", 1, 1), ("RawSigned", 4, 1), ("RawSigned", 4, 0),
                                          ("RawSigned", 8, 1), ("RawSigned", 8, 0), ("RawUnsigned", 12, 1),
                                          ("Mk5B", 4, 1), ("Mk5B", 8, 0)])
def test_unpack_bitexact(hip, oracle, fmt, nch, incr):
    import torch
    say("unpack", fmt, nch, incr)
    count = 8192
    raw = np.random.default_rng(nch * 10 + incr).integers(0, 256, count * max(nch // 4, 1), dtype=np.uint8)
    if nch == 1:
        raw = raw[: count // 4]
    s = hip.derive_settings(FFTpoints=4096, FFTIntegrationTimeSec=1.0, FFToverlapFactor=1, BandwidthHz=8e6,
                            BitsPerSample=2, SourceChannels=nch, SourceFormat=fmt, ChannelOrderIncreasing=bool(incr))
    os_ = oracle_settings(oracle, s)
    d_raw = torch.from_numpy(raw).cuda()
    for ch in range(nch):
        d_out = torch.empty(count, dtype=torch.float32, device="cuda")
        hip.unpack_device(s, d_raw, ch, count, d_out)
        ref = oracle.unpack(os_, raw, count, ch)[:count]
        assert np.array_equal(d_out.cpu().numpy(), ref), (fmt, nch, incr, ch)


# ---------------------------------------------------------------- spectra

SIZES = [
    (4096, "Hann", 5), (20000, "Cosine2", 3), (32000, "Hann", 4), (65536, "Hamming", 3),
    (320000, "Blackman", 2), (1 << 20, "Cosine", 2), (1000000, "None", 2), (3200000, "Cosine2", 2),
    # 3.2M: Blackman with the first column stage fused into the fill; None
    # (no per-column window table) with the column fill
    (3200000, "Blackman", 2), (3200000, "None", 2),
    # 10^6 single-channel with a window: the 8-column row fill
    (1000000, "Cosine", 2),
    # odd row lengths: 2700 = 2 * 30 x 45 (even column count: a self-mirrored
    # middle row with a self-paired bin), 4374 = 2 * 27 x 81, 156250 = 2 * 125 x 625
    (2700, "Hann", 5), (4374, "Cosine2", 4), (156250, "Hamming", 2),
]


@pytest.mark.parametrize("n,win,nf", SIZES)
def test_spectrum_vs_oracle(hip, oracle, n, win, nf):
    say("spectrum", n, win, nf)
    fs = 32e6
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType=win,
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, MaxSourceBufferMB=4096)
    assert s.core_overlapped_ffts == nf
    raw = synth.raw_1ch(n + nf, n * nf, tones=[(0.2, 4.8e6, 0.3)], fs=fs)
    m, auto, _, _, core = run_device(s, [raw])
    os_ = oracle_settings(oracle, s)
    m64, a64, _, _ = oracle.domaths(os_, [raw], 64)
    m32, a32, _, _ = oracle.domaths(os_, [raw], 32)
    assert m == m64 == m32 == 1
    e64 = relerr(auto[0], a64[0])
    e32 = relerr(auto[0], a32[0])
    say("  err vs fp64 %.3e  vs fp32 restatement %.3e" % (e64, e32))
    assert e64 <= TOL and e32 <= TOL
    core.finalize()


def test_multichannel_mk5b_and_pcal(hip, oracle):
    """4-channel Mk5B 3.2M-point geometry at reduced length: 4-channel Mk5B payload, 3.2M points,
    Cosine2, phase-cal comb; all four channels in one pass."""
    n, nf, fs = 3200000, 2, 32e6
    say("3.2M 4ch small", n, nf)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1,
                            WindowType="Cosine2", BandwidthHz=16e6, BitsPerSample=2, SourceChannels=4,
                            SourceFormat="Mk5B", ExtractPCal=True, MaxSourceBufferMB=4096)
    assert s.extract_pcal == 1 and s.pcal_tonebins == 32 and s.core_overlapped_ffts == nf
    rng = np.random.default_rng(42)
    tones = [synth.pcal_comb(fs, fs / 2, amp=0.141, rng=rng) for _ in range(4)]
    raw = synth.raw_multi(7, 4, n * nf, tones_per_ch=tones, fs=fs)
    m, auto, _, pc, core = run_device(s, [raw], channels=[0, 1, 2, 3])
    assert m == 1
    for ch in range(4):
        os_ = oracle_settings(oracle, hip.derive_settings(**dict(s.keys, UseFile1Channel=ch + 1)))
        _, a64, _, p64 = oracle.domaths(os_, [raw], 64)
        e = relerr(auto[ch], a64[0])
        # FP64 fold with the FP32 rotator, as the kernel does
        x = oracle.unpack(os_, raw, n * nf, ch).astype(np.float64)
        acc = np.zeros(64)
        for f in range(nf):
            acc = oracle.pcal_fold64(x[f * n:(f + 1) * n], fs, 10e3, 32, 3200, rot32=True, acc=acc)
        pref = acc[0::2] + 1j * acc[1::2]
        ep = relerr(pc[ch][0], pref)
        ep64 = relerr(pc[ch][0], p64[0][0])
        tones_gpu = core.pcal_tones(ch, 0)
        tones_ref = oracle.pcal_tones64(acc, len(tones_gpu), 1e6, fs)
        et = relerr(tones_gpu, tones_ref)
        say("  ch%d spectrum %.3e pcal %.3e (vs fp64 rotator %.3e) tones %.3e" % (ch, e, ep, ep64, et))
        assert e <= TOL
        assert ep <= 1e-6 and ep64 <= 1e-5 and et <= 1e-6
    core.finalize()


def test_row_stream_with_pcal_matches_one_stream(hip, oracle):
    """Cores with phase-cal whose spectra span several chunks run the row pass
    on a second stream over two Y buffers, beside the next chunk's column pass
    and phase-cal fold.  4-channel Mk5B with phase-cal, 2 spectra of 6 FFTs in
    chunks of 2: spectra and phase-cal bit for bit those of the one-stream
    run (SWSPEC_ROW_STREAM=0), and every channel against the oracle."""
    import os
    n, K, fs = 320000, 6, 32e6
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * K / fs, FFToverlapFactor=1,
                            WindowType="Cosine2", BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=4,
                            SourceFormat="Mk5B", ExtractPCal=True, MaxSourceBufferMB=4096)
    assert s.core_overlapped_ffts == K and s.extract_pcal == 1
    rng = np.random.default_rng(61)
    tones = [synth.pcal_comb(fs, fs / 2, amp=0.15, rng=rng) for _ in range(4)]
    raw = synth.raw_multi(62, 4, n * K * 2, tones_per_ch=tones, fs=fs)
    res = {}
    for mode in ("0", "1"):
        os.environ["SWSPEC_ROW_STREAM"] = mode
        try:
            m, auto, _, pc, core = run_device(s, [raw], channels=[0, 1, 2, 3], ffts_per_chunk=2)
        finally:
            del os.environ["SWSPEC_ROW_STREAM"]
        core.finalize()
        assert m == 2
        res[mode] = (auto, pc)
    for ch in range(4):
        assert np.array_equal(res["0"][0][ch], res["1"][0][ch]), ch
        assert np.array_equal(res["0"][1][ch], res["1"][1][ch]), ch
        os_ = oracle_settings(oracle, hip.derive_settings(**dict(s.keys, UseFile1Channel=ch + 1)))
        m64, a64, _, p64 = oracle.domaths(os_, [raw], 64)
        assert m64 == 2
        e, ep = relerr(res["1"][0][ch], a64[0]), relerr(res["1"][1][ch], p64[0])
        say("  ch%d spectra %.3e pcal %.3e" % (ch, e, ep))
        assert e <= TOL and ep <= 1e-5


@pytest.mark.parametrize("n,win,framed", [(320000, "Blackman", False), (1000000, "None", False),
                                           (1 << 20, "Cosine", False), (3200000, "Hamming", False),
                                           (3200000, "Blackman", True), (320000, "Hann", True),
                                           (3200000, "None", False), (3200000, "Cosine", True),
                                           (1 << 20, "Hann", True),
                                           # 8-column row fill (10^6: columns 625 = 25 x 25)
                                           (1000000, "Hann", False), (1000000, "Blackman", True),
                                           # row fill without the fused first stage: 8 columns of
                                           # 384 = 16 x 24 (16 points per thread), 4 columns of
                                           # 768 = 16 x 48 (16) and of 1200 = 16 x 75 (32)
                                           (307200, "Hann", False), (1572864, "Cosine2", True),
                                           (3072000, "Hamming", False),
                                           # the fused fill in the kernels with the plan's own strides
                                           # (1250 x 1600, 32 points per thread; 1000 x 1000, 16)
                                           (4000000, "Cosine2", False), (2000000, "Hann", True)])
def test_row_fill_windows_vs_oracle(hip, oracle, n, win, framed):
    """The row fill of pass 1 (a thread decodes and windows whole rows of the
    workgroup's 4 or 8 columns from one wide load; one byte per time sample)
    for the window forms it takes: one cosine term (Cosine, Hann, Hamming)
    and the cos(2 theta) term (Blackman), 4- and 8-column plans (3.2M: 4,
    with the radix-10 first column stage fused into the fill; 10^6: 8),
    payload and framed Mk5B; beside them the column fill of the same plans
    (window None has no per-column window table, so it keeps the column
    fill; 320000 and 2^20, two pipelines per workgroup, too).  All four
    channels of a 4-channel Mk5B stream against the FP64 oracle."""
    from swspec_amd import plan_info
    from swspec_amd._lib import FLAG_MK5B_FRAMED
    nf, fs = 2, 32e6
    assert plan_info(n)["B"] in (4, 8)
    say("row fill", n, win, "framed" if framed else "payload")
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType=win,
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=4, SourceFormat="Mk5B",
                            MaxSourceBufferMB=4096)
    # framed input is whole 10000-byte payloads: the bytes past the last FFT
    # (less than one FFT) are read by neither the core nor the oracle
    nbytes = -(-n * nf // 10000) * 10000 if framed else n * nf
    payload = synth.raw_multi(11, 4, nbytes, tones_per_ch=[[(0.2, (3 + c) * 1e6, 0.1 * c)] for c in range(4)],
                              fs=fs)
    data = synth.mk5b_frames(payload) if framed else payload
    m, auto, _, _, core = run_device(s, [data], channels=[0, 1, 2, 3], flags=FLAG_MK5B_FRAMED if framed else 0)
    core.finalize()
    assert m == 1
    for ch in range(4):
        os_ = oracle_settings(oracle, hip.derive_settings(**dict(s.keys, UseFile1Channel=ch + 1)))
        _, a64, _, _ = oracle.domaths(os_, [payload], 64)
        e = relerr(auto[ch], a64[0])
        say("  ch%d spectrum %.3e" % (ch, e))
        assert e <= TOL


@pytest.mark.parametrize("nch,channels,framed", [(8, list(range(8)), False), (4, [1, 3], False),
                                                 (4, [0, 1, 2, 3], True), (8, [4, 6], True),
                                                 (8, [0, 2], False), (8, [4, 5, 6, 7], True)])
def test_channel_subsets_vs_oracle(hip, oracle, nch, channels, framed):
    """Several channels of one stream in one core (jobs sharing the stream's
    bytes in the same launches): all 8 channels of 8-channel Mk5B (bytes 0
    and 1 of each time sample), a pair of channels, and framed input; every
    channel's spectrum and phase-cal against the oracle, and within 1e-6
    (phase-cal bit for bit) of the channel run alone."""
    from swspec_amd._lib import FLAG_MK5B_FRAMED
    n, nf, fs = 320000, 3, 32e6
    say("channel groups", nch, channels, framed)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1,
                            WindowType="Cosine2", BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=nch,
                            SourceFormat="Mk5B", ExtractPCal=True, MaxSourceBufferMB=4096)
    rng = np.random.default_rng(50 + nch + len(channels))
    tones = [synth.pcal_comb(fs, fs / 2, amp=0.15, rng=rng) for _ in range(nch)]
    payload = synth.raw_multi(5, nch, n * nf, tones_per_ch=tones, fs=fs)
    data = synth.mk5b_frames(payload) if framed else payload
    flags = FLAG_MK5B_FRAMED if framed else 0
    m, auto, _, pc, core = run_device(s, [data], channels=channels, flags=flags)
    core.finalize()
    assert m == 1
    for j, ch in enumerate(channels):
        sc = hip.derive_settings(**dict(s.keys, UseFile1Channel=ch + 1))
        _, a64, _, p64 = oracle.domaths(oracle_settings(oracle, sc), [payload], 64)
        e, ep = relerr(auto[j], a64[0]), relerr(pc[j], p64[0])
        m1, a1, _, p1, c1 = run_device(sc, [data], flags=flags)
        c1.finalize()
        say("  ch%d spectrum %.3e pcal %.3e" % (ch, e, ep))
        assert e <= TOL and ep <= 1e-5
        assert relerr(auto[j], a1[0]) <= 1e-6 and np.array_equal(pc[j], p1[0])


@pytest.mark.parametrize("fmt,nch,incr,bw,n", [
    ("RawSigned", 1, 1, 16e6, 32000),     # fast fold, 4 samples per byte
    ("RawSigned", 8, 0, 16e6, 6400),      # fast fold, 2 bytes per time sample
    ("Mk5B", 8, 1, 32e6, 12800),          # 64 Ms/s: 64 tone bins, 6400-sample rotator
    ("RawSigned", 4, 1, 9e6, 18000),      # 18 tone bins: generic fold kernel
    # 1600 / 3200 groups of 4 rotator positions: 2 / 4 groups per thread
    ("RawSigned", 1, 1, 32e6, 64000),
    ("RawSigned", 1, 1, 64e6, 128000),
    ("Mk5B", 4, 1, 32e6, 64000),
    ("Mk5B", 4, 1, 64e6, 128000),
    ("Mk5B", 8, 1, 64e6, 128000),
    # 16 channels: 4 bytes per time sample, 1 / 2 / 4 groups per thread
    ("RawSigned", 16, 1, 16e6, 32000),
    ("RawSigned", 16, 1, 32e6, 64000),
    ("RawUnsigned", 16, 0, 64e6, 128000),
])
def test_pcal_formats(hip, oracle, fmt, nch, incr, bw, n):
    say("pcal", fmt, nch, incr, bw, n)
    fs = 2 * bw
    nf = 3
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=bw, BitsPerSample=2, SourceChannels=nch, SourceFormat=fmt,
                            ChannelOrderIncreasing=bool(incr), ExtractPCal=True, MaxSourceBufferMB=4096,
                            UseFile1Channel=nch)
    assert s.extract_pcal == 1
    rng = np.random.default_rng(int(bw) % 1000 + n)
    comb = synth.pcal_comb(fs, bw, amp=0.2, rng=rng)
    if nch == 1:
        raw = synth.raw_1ch(n, n * nf, tones=comb, fs=fs)
    else:
        raw = synth.raw_multi(n, nch, n * nf, tones_per_ch=[comb] * nch, fs=fs, increasing=bool(incr))
    m, auto, _, pc, core = run_device(s, [raw])
    os_ = oracle_settings(oracle, s)
    _, a64, _, p64 = oracle.domaths(os_, [raw], 64)
    x = oracle.unpack(os_, raw, n * nf, nch - 1).astype(np.float64)
    acc = np.zeros(2 * s.pcal_tonebins)
    for f in range(nf):
        acc = oracle.pcal_fold64(x[f * n:(f + 1) * n], fs, 10e3, s.pcal_tonebins, s.pcal_rotatorlen, rot32=True,
                                 acc=acc)
    ep = relerr(pc[0][0], acc[0::2] + 1j * acc[1::2])
    tg = core.pcal_tones(0, 0)
    tr = oracle.pcal_tones64(acc, len(tg), 1e6, fs)
    say("  T=%d L=%d pcal %.3e vs-fp64 %.3e tones %.3e spectrum %.3e" % (
        s.pcal_tonebins, s.pcal_rotatorlen, ep, relerr(pc[0][0], p64[0][0]), relerr(tg, tr), relerr(auto[0], a64[0])))
    assert m == 1 and ep <= 1e-6 and relerr(tg, tr) <= 1e-6 and relerr(auto[0], a64[0]) <= TOL
    core.finalize()


def test_xpol_vs_oracle(hip, oracle):
    n, nf, fs = 1 << 16, 4, 64e6
    say("xpol", n, nf)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, DoCrossPolarization=True,
                            MaxSourceBufferMB=4096, num_sources=2)
    assert s.num_xpols == 1
    rng = np.random.default_rng(3)
    a = rng.standard_normal(n * nf)
    b = 0.2 * a + np.sqrt(1 - 0.04) * rng.standard_normal(n * nf)
    ra, rb = synth.pack_1ch(synth.quantize(a)), synth.pack_1ch(synth.quantize(b))
    m, auto, xp, _, core = run_device(s, [ra, rb])
    os_ = oracle_settings(oracle, s)
    _, a64, x64, _ = oracle.domaths(os_, [ra, rb], 64)
    _, a32, x32, _ = oracle.domaths(os_, [ra, rb], 32)
    assert m == 1
    ex = relerr(xp[0], x64[0])
    say("  auto %.3e %.3e xpol %.3e (fp32 restatement %.3e)" % (relerr(auto[0], a64[0]), relerr(auto[1], a64[1]),
                                                               ex, relerr(xp[0], x32[0])))
    assert relerr(auto[0], a64[0]) <= TOL and relerr(auto[1], a64[1]) <= TOL
    assert ex <= TOL and relerr(xp[0], x32[0]) <= TOL
    assert xp[0][0, 0].imag == 0 and xp[0][0, -1].imag == 0
    core.finalize()


def test_xpol_3p2m_constant_stride_kernels(hip, oracle):
    """Cross-pol pair at 3.2M points (single-channel 2-bit, 32 Ms/s, Hann):
    the plan is the one with compile-time strides (columns 1250 = 10*25*5,
    rows 1280 = 5*16*16 with the padded stage), so pass 1 runs the
    constant-stride fused-fill kernel and pass 2 its two-source constant-
    stride form; X, Y, XY* vs the FP64 oracle (TaskCoreIPP.cpp:438-498)."""
    n, nf, fs = 3200000, 2, 32e6
    say("xpol 3.2M", n, nf)
    pi = hip.plan_info(n)
    assert (pi["N1"], pi["N2"], pi["B"], pi["p1_radix0"], pi["p2_pad"]) == (1250, 1280, 4, 10, 5), pi
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, DoCrossPolarization=True,
                            MaxSourceBufferMB=4096, num_sources=2)
    assert s.num_xpols == 1 and s.core_overlapped_ffts == nf
    rng = np.random.default_rng(33)
    a = rng.standard_normal(n * nf)
    b = 0.3 * a + np.sqrt(1 - 0.09) * rng.standard_normal(n * nf)
    ra, rb = synth.pack_1ch(synth.quantize(a)), synth.pack_1ch(synth.quantize(b))
    m, auto, xp, _, core = run_device(s, [ra, rb])
    os_ = oracle_settings(oracle, s)
    _, a64, x64, _ = oracle.domaths(os_, [ra, rb], 64)
    assert m == 1
    ea, eb, ex = relerr(auto[0], a64[0]), relerr(auto[1], a64[1]), relerr(xp[0], x64[0])
    say("  auto %.3e %.3e xpol %.3e" % (ea, eb, ex))
    assert ea <= TOL and eb <= TOL and ex <= TOL
    assert xp[0][0, 0].imag == 0 and xp[0][0, -1].imag == 0
    core.finalize()


@pytest.mark.parametrize("n,knobs", [
    (65536, {"SWSPEC_COMPOSITE": "0"}),
    (1000000, {"SWSPEC_COMPOSITE": "0"}),
    (65536, {"SWSPEC_P2_MAXPTS": "32"}),
    (1000000, {"SWSPEC_P2_MAXPTS": "32"}),
    (1000000, {"SWSPEC_P2_MAXPTS": "32", "SWSPEC_COMPOSITE": "0"}),
    (1000000, {"SWSPEC_P1_MAXPTS": "16"}),
    (4096, {"SWSPEC_COMPOSITE": "0"}),   # single-pass short kernel, two sources, plain radices
    (16384, {"SWSPEC_COMPOSITE": "0"}),  # the same with 32 points per thread
    (1000000, {"SWSPEC_COMPOSITE": "0", "SWSPEC_P1_MAXPTS": "32"}),   # 8-column row fill, plain, 32 points
    (524288, {"SWSPEC_COMPOSITE": "0"}),   # rows 512: plain first radix 8
    (294912, {"SWSPEC_COMPOSITE": "0"}),   # rows 384: plain first radix 3
])
def test_xpol_plan_variants(hip, oracle, n, knobs):
    """Cross-pol pairs on the plan variants the autotuner can choose (plain
    radices only, 32 row points per thread, 16 column points per thread):
    the two-source row kernels of those plans, X, Y and XY* vs the FP64
    oracle.  The knobs are read when the core builds its plan (prepare)."""
    import os
    nf, fs = 2, 64e6
    say("xpol plan", n, knobs)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, DoCrossPolarization=True,
                            MaxSourceBufferMB=4096, num_sources=2)
    rng = np.random.default_rng(n % 1000 + len(knobs))
    a = rng.standard_normal(n * nf)
    b = 0.25 * a + np.sqrt(1 - 0.0625) * rng.standard_normal(n * nf)
    ra, rb = synth.pack_1ch(synth.quantize(a)), synth.pack_1ch(synth.quantize(b))
    os.environ.update(knobs)
    try:
        m, auto, xp, _, core = run_device(s, [ra, rb])
    finally:
        for k in knobs:
            del os.environ[k]
    core.finalize()
    os_ = oracle_settings(oracle, s)
    _, a64, x64, _ = oracle.domaths(os_, [ra, rb], 64)
    assert m == 1
    ea, eb, ex = relerr(auto[0], a64[0]), relerr(auto[1], a64[1]), relerr(xp[0], x64[0])
    say("  auto %.3e %.3e xpol %.3e" % (ea, eb, ex))
    assert ea <= TOL and eb <= TOL and ex <= TOL


def test_short_plain_radices_single_source(hip, oracle):
    """Single-pass short FFT with 32 points per thread on plain radices (the
    autotuner's cx_mask 0 at 16384 points), one source, vs the FP64 oracle."""
    import os
    n, nf, fs = 16384, 6, 32e6
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType="Hamming",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, MaxSourceBufferMB=4096)
    raw = synth.raw_1ch(n + 3, n * nf, tones=[(0.2, 4.8e6, 0.3)], fs=fs)
    os.environ["SWSPEC_COMPOSITE"] = "0"
    try:
        info = hip.plan_info(n)
        m, auto, _, _, core = run_device(s, [raw])
    finally:
        del os.environ["SWSPEC_COMPOSITE"]
    core.finalize()
    assert info["single"] == 1 and info["p1_maxpts"] == 32 and info["p1_cx"] == 0, info
    os_ = oracle_settings(oracle, s)
    _, a64, _, _ = oracle.domaths(os_, [raw], 64)
    assert m == 1
    e = relerr(auto[0], a64[0])
    say("  short plain 32: err vs fp64 %.3e" % e)
    assert e <= TOL


def test_schedule_overlap_partial_multispectrum(hip, oracle):
    """Overlap factor 2, several spectra per buffer, trailing partial spectrum."""
    n, fs = 6400, 32e6     # N % rotatorlen (3200) == 0 keeps phase-cal on
    say("schedule overlap")
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * 3 / fs, FFToverlapFactor=2, WindowType="Cosine",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, MaxSourceBufferMB=4096,
                            ExtractPCal=True, PCalOffsetHz=10e3)
    K = s.core_overlapped_ffts
    assert K == 5
    raw = synth.raw_1ch(1, n * 3 * 3 + n // 2 * 3, fs=fs)   # 3 full spectra + partial
    m, auto, _, pc, core = run_device(s, [raw], ffts_per_chunk=2)
    os_ = oracle_settings(oracle, s)
    m64, a64, _, p64 = oracle.domaths(os_, [raw], 64)
    assert m == m64 == 4
    assert relerr(auto[0], a64[0]) <= TOL
    say("  pcal err %.3e" % relerr(pc[0], p64[0]))
    assert relerr(pc[0], p64[0]) <= 1e-5
    core.finalize()


def test_host_mode_matches_device_mode(hip, oracle):
    n, nf, fs = 32000, 6, 32e6
    say("host mode")
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, MaxSourceBufferMB=4096)
    raw = synth.raw_1ch(9, n * nf * 2, fs=fs)   # two spectra in one buffer
    m, auto, _, _, core = run_device(s, [raw])
    core.finalize()
    c2 = hip.TaskCoreHIP(1)
    c2.prepare(s)
    inb = hip.Buffer(len(raw), pinned=True)
    inb.getData()[:] = raw
    inb.setLength(len(raw))
    out = hip.Buffer(2 * s.fft_bytes_ssb)
    c2.run([inb], [out], [], [])
    k = c2.join()
    assert k == m == 2
    assert out.getLength() == 2 * s.fft_bytes_ssb
    host = out.getData().view(np.float32).reshape(2, -1)
    assert np.array_equal(host, auto[0])
    # combineResults adds the first spectrum into a combined buffer (TaskCoreIPP.cpp:311-349)
    comb = hip.Buffer(s.fft_bytes_ssb)
    c2.resetBuffer(comb)
    c2.combineResults([comb], [])
    assert np.array_equal(comb.getData().view(np.float32), auto[0][0])
    c2.finalize()


@pytest.mark.parametrize("seg_mb", ["1", "0", None])
def test_host_segmented_input_matches_device_mode(hip, oracle, seg_mb):
    """Host Buffers in pinned memory, copied in SWSPEC_H2D_SEG_MB segments on
    the copy stream while earlier chunks compute (0: one copy; None: the
    default 32 MB): spectra, cross spectrum and phase-cal bit-identical to the
    same run on device-resident input (the chunking, and so the summation
    order, is the same).  4-channel Mk5B with phase-cal over 8 chunks, and a
    2-source cross-pol pair over 4 chunks."""
    import os
    say("host segmented input", seg_mb)
    old = os.environ.pop("SWSPEC_H2D_SEG_MB", None)
    if seg_mb is not None:
        os.environ["SWSPEC_H2D_SEG_MB"] = seg_mb
    try:
        fs = 32e6
        cases = []
        n, nf = 320000, 30
        s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1,
                                WindowType="Cosine2", BandwidthHz=16e6, BitsPerSample=2, SourceChannels=4,
                                SourceFormat="Mk5B", ExtractPCal=True, MaxSourceBufferMB=4096)
        rng = np.random.default_rng(8)
        tones = [synth.pcal_comb(fs, fs / 2, amp=0.141, rng=rng) for _ in range(4)]
        cases.append((s, [synth.raw_multi(3, 4, n * nf, tones_per_ch=tones, fs=fs)], [0, 1, 2, 3]))
        n, nf = 1 << 20, 8
        s2 = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / 64e6, FFToverlapFactor=1,
                                 WindowType="Hann", BandwidthHz=32e6, BitsPerSample=2, SourceChannels=1,
                                 DoCrossPolarization=True, MaxSourceBufferMB=4096, num_sources=2)
        cases.append((s2, [synth.raw_1ch(31, n * nf), synth.raw_1ch(32, n * nf)], None))
        for st, raws, chans in cases:
            m, auto, xp, pc, core = run_device(st, raws, channels=chans, ffts_per_chunk=2 if chans is None else 4)
            core.finalize()
            c2 = hip.TaskCoreHIP(1)
            c2.prepare(st, channels=chans, ffts_per_chunk=2 if chans is None else 4)
            ins = []
            for r in raws:
                b = hip.Buffer(len(r), pinned=True)
                b.getData()[:] = r
                b.setLength(len(r))
                ins.append(b)
            nj = c2.njobs
            outs = [hip.Buffer(st.fft_bytes_ssb) for _ in range(nj)]
            xb = [hip.Buffer(2 * st.fft_bytes_ssb)] if st.num_xpols else []
            T = st.pcal_tonebins if st.extract_pcal else 0
            pb = [hip.Buffer(8 * T) for _ in range(nj)] if T else []
            c2.run(ins, outs, xb, pb)
            assert c2.join() == m == 1
            for j in range(nj):
                assert np.array_equal(outs[j].getData().view(np.float32), auto[j][0]), (seg_mb, j)
                if T:
                    v = pb[j].getData().view(np.float32)
                    assert np.array_equal(v[0::2] + 1j * v[1::2], pc[j][0].astype(np.complex64)), (seg_mb, j)
            if xb:
                v = xb[0].getData().view(np.float32)
                assert np.array_equal(v[0::2] + 1j * v[1::2], xp[0][0].astype(np.complex64)), seg_mb
            c2.finalize()
    finally:
        os.environ.pop("SWSPEC_H2D_SEG_MB", None)
        if old is not None:
            os.environ["SWSPEC_H2D_SEG_MB"] = old


def test_combine_device_matches_host_combine(hip):
    """combineResults on device arrays (auto, cross and phase-cal outputs of
    two sources, two runs added up) equals the host combineResults of the
    same runs bit for bit (TaskCoreIPP.cpp:311-349)."""
    import torch
    n, fs = 32000, 32e6
    say("combine_device")
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * 2 / fs, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, DoCrossPolarization=True,
                            ExtractPCal=True, PCalOffsetHz=10e3, MaxSourceBufferMB=4096, num_sources=2)
    assert s.num_xpols == 1 and s.extract_pcal == 1
    ssb, T = s.fft_ssb_points, s.pcal_tonebins
    runs = [[synth.raw_1ch(40 + 2 * r + k, n * 2, fs=fs) for k in range(2)] for r in range(2)]
    # host mode: run, join, combineResults twice
    hc = hip.TaskCoreHIP(1)
    hc.prepare(s)
    h_comb = [hip.Buffer(4 * ssb) for _ in range(2)] + [hip.Buffer(8 * ssb)]
    h_pc = [hip.Buffer(8 * T) for _ in range(2)]
    for b in h_comb + h_pc:
        hc.resetBuffer(b)
    for raws in runs:
        ins = []
        for r in raws:
            b = hip.Buffer(len(r), pinned=True)
            b.getData()[:] = r
            b.setLength(len(r))
            ins.append(b)
        outs = [hip.Buffer(4 * ssb) for _ in range(2)]
        xb = [hip.Buffer(8 * ssb)]
        pb = [hip.Buffer(8 * T) for _ in range(2)]
        hc.run(ins, outs, xb, pb)
        assert hc.join() == 1
        hc.combineResults(h_comb, h_pc)
    hc.finalize()
    # device mode: run_device, join, combine_device twice
    dev = torch.device("cuda:0")
    dc = hip.TaskCoreHIP(0)
    dc.prepare(s)
    d_comb = [torch.full((ssb,), 7.0, device=dev) for _ in range(2)] + [torch.full((2 * ssb,), 7.0, device=dev)]
    d_pc = [torch.full((2 * T,), 7.0, device=dev) for _ in range(2)]
    for t in d_comb + d_pc:
        hip.TaskCoreHIP.reset_device(t)
    assert all(float(t.abs().sum()) == 0 for t in d_comb + d_pc)
    for raws in runs:
        d_in = [torch.from_numpy(r).to(dev) for r in raws]
        d_auto = [torch.full((ssb,), float("nan"), device=dev) for _ in range(2)]
        d_x = [torch.full((2 * ssb,), float("nan"), device=dev)]
        d_p = [torch.full((2 * T,), float("nan"), device=dev) for _ in range(2)]
        dc.run_device(d_in, [len(r) for r in raws], d_auto, d_x, d_p)
        assert dc.join() == 1
        ol, pl = dc.combine_device(d_comb, d_pc)
        assert ol == [4 * ssb, 4 * ssb, 4 * ssb]   # cross length float-sized, as the reference sets it
        assert pl == [8 * T, 8 * T]
        # the core's outputs are cleared once added
        assert all(float(t.abs().sum()) == 0 for t in d_auto + d_x + d_p)
    for k in range(3):
        h = h_comb[k].getData().view(np.float32)[: d_comb[k].numel()]
        assert np.array_equal(d_comb[k].cpu().numpy(), h), k
    for k in range(2):
        assert np.array_equal(d_pc[k].cpu().numpy(), h_pc[k].getData().view(np.float32)[: 2 * T]), k
    assert float(d_comb[0].sum()) > 0
    # combine before any run-in-progress is joined is refused
    d_in = [torch.from_numpy(r).to(dev) for r in runs[0]]
    d_auto = [torch.empty(ssb, device=dev) for _ in range(2)]
    dc.run_device(d_in, [len(r) for r in runs[0]], d_auto, [torch.empty(2 * ssb, device=dev)],
                  [torch.empty(2 * T, device=dev) for _ in range(2)])
    with pytest.raises(RuntimeError):
        dc.combine_device(d_comb, d_pc)
    assert dc.join() == 1
    dc.finalize()


def test_parseval_full_3p2m_length(hip):
    """Size-independent property at the full 3.2M-point length with many FFTs:
    sum_k c_k P[k] = N * mean_f sum_n (w x)^2 (c = 1 at DC/Nyquist, 2 otherwise)."""
    import torch
    n, nf, fs = 3200000, 12, 32e6
    say("parseval 3.2M x", nf)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1,
                            WindowType="Cosine2", BandwidthHz=16e6, BitsPerSample=2, SourceChannels=4,
                            SourceFormat="Mk5B", MaxSourceBufferMB=4096)
    raw = np.random.default_rng(77).integers(0, 256, n * nf, dtype=np.uint8)
    m, auto, _, _, core = run_device(s, [raw], channels=[2])
    P = auto[0][0].astype(np.float64)
    lhs = P[0] + P[-1] + 2 * P[1:-1].sum()
    lev = torch.tensor(synth.LEVELS, dtype=torch.float64, device="cuda")
    x = lev[((torch.from_numpy(raw).cuda().long() >> 4) & 3)].reshape(nf, n)
    i = torch.arange(n, dtype=torch.float64, device="cuda")
    w = torch.sin(np.pi * (2 * i + 1) / (2 * n)) ** 2
    rhs = n * float(((x * w) ** 2).sum(dim=1).mean())
    say("  parseval rel %.3e" % (abs(lhs - rhs) / rhs))
    assert abs(lhs - rhs) <= 1e-5 * rhs
    core.finalize()


# ---------------------------------------------------------------- VLBA (mark5access decode)

def _vlba_bytes(rng, n, fills):
    """Random 8-track bytes with Mark5 fill-pattern zones [a, b)."""
    raw = rng.integers(0, 256, n, dtype=np.uint8)
    raw[::4] = np.where(raw[::4] == 0x44, 0x45, raw[::4])
    pat = np.frombuffer(bytes([0x44, 0x33, 0x22, 0x11]), dtype=np.uint8)
    for a, b in fills:
        for o in range(a, b, 4):
            raw[o:o + 4] = pat
    return raw


VLBA_FILLS = [(12000, 20000), (40000 + 40128, 40128 + 40128 + 8),   # tail of a payload; the 128-byte chunk tail
              (2 * 40128, 2 * 40128 + 4096), (3 * 40128 + 20000 + 56, 3 * 40128 + 40000)]


@pytest.mark.parametrize("ch", [0, 1, 2, 3])
def test_unpack_vlba_bitexact(hip, oracle, ch):
    """VLBA1_1-128-4-2 decode (modulation, blank zones, per-chunk restart)
    bit-exact against the oracle's restatement of VLBAUnpacker."""
    import torch
    say("unpack vlba", ch)
    count = 4 * 40128 + 5000
    raw = _vlba_bytes(np.random.default_rng(11 + ch), count, VLBA_FILLS)
    s = hip.derive_settings(FFTpoints=40000, FFTIntegrationTimeSec=0.01, FFToverlapFactor=1, BandwidthHz=8e6,
                            BitsPerSample=2, SourceChannels=4, SourceFormat="VLBA1_1-128-4-2")
    d_raw = torch.from_numpy(raw).cuda()
    d_out = torch.empty(count, dtype=torch.float32, device="cuda")
    hip.unpack_device(s, d_raw, ch, count, d_out)
    ref = oracle.vlba_unpack(raw, count, ch, avail=count)
    got = d_out.cpu().numpy()
    assert np.array_equal(got, ref), int(np.argmax(got != ref))
    assert np.count_nonzero(ref == 0) > 0   # the fill zones were blanked


def test_vlba_spectrum_and_pcal_vs_oracle(hip, oracle):
    """doMaths over VLBA bytes with overlapping FFTs (every FFT restarts the
    chunking), phase-cal on (generic fold) and blank zones inside FFTs."""
    n, nf, fs = 96000, 3, 32e6     # 96000 = 30 rotator periods of 3200
    say("vlba spectrum", n, nf)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=2, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=4, SourceFormat="VLBA1_1-256-4-2",
                            ExtractPCal=True, MaxSourceBufferMB=4096, UseFile1Channel=3)
    assert s.extract_pcal == 1
    nbytes = n * (nf + 2)
    raw = _vlba_bytes(np.random.default_rng(21), nbytes, [(30000, 36000), (100000, 140000), (200008, 210000)])
    m, auto, _, pc, core = run_device(s, [raw])
    os_ = oracle_settings(oracle, s)
    m64, a64, _, p64 = oracle.domaths(os_, [raw], 64)
    assert m == m64
    e = relerr(auto[0], a64[0])
    ep = relerr(pc[0], p64[0])
    say("  spectra %d err %.3e  pcal err %.3e" % (m, e, ep))
    assert e <= TOL and ep <= 1e-5
    core.finalize()


VLBA_MODES = [(1, 2), (2, 2), (4, 2), (1, 1), (2, 1), (4, 1)]   # (fanout, nbit): one byte per time sample


@pytest.mark.parametrize("fanout,nbit", VLBA_MODES)
def test_unpack_vlba_modes_bitexact(hip, oracle, fanout, nbit):
    """Every one-byte-per-sample VLBA mode (VLBA1_<fanout>-*-4-2 and -8-1 at
    fanout 1, 2, 4: 8, 16, 32 tracks), every channel, bit-exact against the
    oracle over two chunks and a part: fill at the end of a payload (its last
    blank zone), fill from a chunk's tail across the chunk boundary, and a
    fill inside a zone (the blanker only tests zone ends: decoded as data)."""
    import torch
    nch = 4 if nbit == 2 else 8
    T = 8 * fanout
    P, CH = 2500 * T, 5016 * T
    count = 2 * CH + 3 * P // 2
    say("unpack vlba fanout %d nbit %d" % (fanout, nbit))
    fills = [(P - 5000, P), (2 * P + 8, CH + 4096), (CH + P + 33000, CH + P + 34000)]
    raw = _vlba_bytes(np.random.default_rng(70 + 3 * fanout + nbit), count, fills)
    s = hip.derive_settings(FFTpoints=40000, FFTIntegrationTimeSec=0.01, FFToverlapFactor=1, BandwidthHz=8e6,
                            BitsPerSample=nbit, SourceChannels=nch,
                            SourceFormat="VLBA1_%d-256-%d-%d" % (fanout, nch, nbit))
    d_raw = torch.from_numpy(raw).cuda()
    for ch in range(nch):
        d_out = torch.empty(count, dtype=torch.float32, device="cuda")
        hip.unpack_device(s, d_raw, ch, count, d_out)
        ref = oracle.vlba_unpack(raw, count, ch, avail=count, fanout=fanout, nbit=nbit)
        got = d_out.cpu().numpy()
        assert np.array_equal(got, ref), (ch, int(np.argmax(got != ref)))
        assert np.count_nonzero(ref == 0) >= 5000


@pytest.mark.parametrize("fanout,nbit", [(2, 2), (4, 1)])
def test_vlba_modes_spectrum_and_pcal_vs_oracle(hip, oracle, fanout, nbit):
    """doMaths over 16-track 2-bit and 32-track 1-bit VLBA bytes: overlapping
    FFTs, phase-cal (generic fold), blank zones inside FFTs."""
    nch = 4 if nbit == 2 else 8
    n, nf, fs = 192000, 3, 32e6      # 60 rotator periods of 3200
    say("vlba spectrum fanout %d nbit %d" % (fanout, nbit))
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=2, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=nbit, SourceChannels=nch,
                            SourceFormat="VLBA1_%d-512-%d-%d" % (fanout, nch, nbit), ExtractPCal=True,
                            MaxSourceBufferMB=4096, UseFile1Channel=nch - 1)
    assert s.extract_pcal == 1
    nbytes = n * (nf + 2)
    raw = _vlba_bytes(np.random.default_rng(23 + fanout), nbytes,
                      [(30000, 42000), (100000, 180000), (250008, 290000)])
    m, auto, _, pc, core = run_device(s, [raw])
    os_ = oracle_settings(oracle, s)
    m64, a64, _, p64 = oracle.domaths(os_, [raw], 64)
    assert m == m64
    e = relerr(auto[0], a64[0])
    ep = relerr(pc[0], p64[0])
    say("  spectra %d err %.3e  pcal err %.3e" % (m, e, ep))
    assert e <= TOL and ep <= 1e-5
    core.finalize()


# ---------------------------------------------------------------- Mark5B (mark5access bitstreams)

MARK5B_CONFIGS = [(2, 1), (4, 1), (8, 1), (16, 1), (32, 1), (2, 2), (4, 2), (8, 2), (16, 2)]   # (nchan, nbit)


@pytest.mark.parametrize("nchan,nbit", MARK5B_CONFIGS)
def test_unpack_mark5b_bitexact(hip, oracle, nchan, nbit):
    """Mark5BUnpacker over mark5access (Mark5B-<Mbps>-<nchan>-<nbit>), every
    bitstream count the reference admits, bit-exact against the oracle over
    four 10000-byte payloads, with a fill zone at the end of one payload and
    over the start of the next (blanked from the data side in each zone)."""
    import torch
    B = nchan * nbit
    count = (4 * 10000 * 8) // B + 8 * 16
    nbytes = count * B // 8
    say("unpack mark5b %d ch x %d bit" % (nchan, nbit))
    raw = _vlba_bytes(np.random.default_rng(90 + B + nbit), nbytes, [(17000, 20000), (30000, 33000)])
    s = hip.derive_settings(FFTpoints=40000, FFTIntegrationTimeSec=0.01, FFToverlapFactor=1, BandwidthHz=8e6,
                            BitsPerSample=nbit, SourceChannels=nchan,
                            SourceFormat="Mark5B-%d-%d-%d" % (256 * B // 8, nchan, nbit))
    d_raw = torch.from_numpy(raw).cuda()
    for ch in sorted({0, 1, nchan // 2, nchan - 1}):
        d_out = torch.empty(count, dtype=torch.float32, device="cuda")
        hip.unpack_device(s, d_raw, ch, count, d_out)
        ref = oracle.mark5b_unpack(raw, count, ch, nchan, nbit, avail=nbytes)
        got = d_out.cpu().numpy()
        assert np.array_equal(got, ref), (ch, int(np.argmax(got != ref)))
        assert np.count_nonzero(ref == 0) > 0


@pytest.mark.parametrize("nchan,nbit", [(4, 2), (16, 1)])
def test_mark5b_spectrum_and_pcal_vs_oracle(hip, oracle, nchan, nbit):
    """doMaths over Mark5B bitstream bytes with overlapping FFTs, phase-cal
    (generic fold) and blank zones inside FFTs."""
    n, nf, fs = 64000, 3, 32e6
    B = nchan * nbit
    say("mark5b spectrum %d ch x %d bit" % (nchan, nbit))
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=2, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=nbit, SourceChannels=nchan,
                            SourceFormat="Mark5B-%d-%d-%d" % (B * 32, nchan, nbit), ExtractPCal=True,
                            MaxSourceBufferMB=4096, UseFile1Channel=nchan - 1)
    assert s.extract_pcal == 1 and s.raw_fullfft_bytes == n * B // 8
    nbytes = n * B // 8 * (nf + 2)
    raw = _vlba_bytes(np.random.default_rng(31 + B), nbytes, [(5000, 10000), (26000, 31000), (60008, 70000)])
    m, auto, _, pc, core = run_device(s, [raw])
    os_ = oracle_settings(oracle, s)
    m64, a64, _, p64 = oracle.domaths(os_, [raw], 64)
    assert m == m64
    e = relerr(auto[0], a64[0])
    ep = relerr(pc[0], p64[0])
    say("  spectra %d err %.3e  pcal err %.3e" % (m, e, ep))
    assert e <= TOL and ep <= 1e-5
    core.finalize()


# ---------------------------------------------------------------- Mk5B frames on the device

def test_unpack_mk5b_framed_bitexact(hip, oracle):
    """Whole Mk5B frames in, headers skipped on the device: equals the oracle's
    FileSource header strip followed by the Mk5B decoder."""
    import torch
    from swspec_amd._lib import FLAG_MK5B_FRAMED
    say("unpack mk5b framed")
    for nch in (4, 8):
        payload = np.random.default_rng(31 + nch).integers(0, 256, 7 * 10000, dtype=np.uint8)
        frames = synth.mk5b_frames(payload)
        count = len(payload) // (nch // 4)
        s = hip.derive_settings(FFTpoints=4000, FFTIntegrationTimeSec=0.01, FFToverlapFactor=1, BandwidthHz=8e6,
                                BitsPerSample=2, SourceChannels=nch, SourceFormat="Mk5B")
        os_ = oracle_settings(oracle, s)
        stripped = oracle.mk5b_strip(frames)
        d_raw = torch.from_numpy(frames).cuda()
        for ch in range(nch):
            d_out = torch.empty(count, dtype=torch.float32, device="cuda")
            hip.unpack_device(s, d_raw, ch, count, d_out, flags=FLAG_MK5B_FRAMED)
            ref = oracle.unpack(os_, stripped, count, ch)[:count]
            assert np.array_equal(d_out.cpu().numpy(), ref), (nch, ch)


@pytest.mark.parametrize("nch,fs", [(4, 32e6), (8, 32e6), (4, 64e6), (8, 64e6), (4, 128e6), (8, 128e6)])
def test_mk5b_framed_run_matches_stripped(hip, oracle, nch, fs):
    """A core fed the on-disk frames gives the spectra of the core fed the
    stripped payload bit for bit (same decoded samples, same FFT), and both
    match the oracle; phase-cal bit for bit too (the same fold reads each
    payload word -- 4 channels -- or 8-byte group -- 8 channels -- through
    its frame address).  32 / 64 / 128 Ms/s: rotators of 3200 / 6400 / 12800
    samples, 1 / 2 / 4 groups of 4 positions per thread."""
    from swspec_amd._lib import FLAG_MK5B_FRAMED
    n, nf = 320000, 3
    say("mk5b framed run", n, nf, nch)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=2, WindowType="Cosine2",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=nch, SourceFormat="Mk5B",
                            ExtractPCal=True, MaxSourceBufferMB=4096, UseFile1Channel=nch - 2)
    rng = np.random.default_rng(33)
    comb = synth.pcal_comb(fs, fs / 2, amp=0.2, rng=rng)
    payload = synth.raw_multi(7, nch, n * (nf + 2), tones_per_ch=[comb] * nch, fs=fs)
    payload = payload[: (len(payload) // 10000) * 10000]
    frames = synth.mk5b_frames(payload)
    m1, a1, _, p1, c1 = run_device(s, [payload])
    c1.finalize()
    from swspec_amd import TaskCoreHIP
    import torch
    core = TaskCoreHIP(0)
    core.prepare(s, flags=FLAG_MK5B_FRAMED)
    ssb, T = s.fft_ssb_points, s.pcal_tonebins
    d_in = torch.from_numpy(frames).cuda()
    d_a = torch.full((4 * ssb,), float("nan"), device="cuda")
    d_p = torch.full((8 * T,), float("nan"), device="cuda")
    core.run_device([d_in], [len(frames)], [d_a], None, [d_p])
    m2 = core.join()
    core.finalize()
    a2 = d_a.cpu().numpy().reshape(4, ssb)[:m2]
    pv = d_p.cpu().numpy()
    p2 = (pv[0::2] + 1j * pv[1::2]).reshape(4, T)[:m2]
    assert m1 == m2 >= 1
    assert np.array_equal(a1[0], a2)
    ep = relerr(p2, p1[0])
    os_ = oracle_settings(oracle, s)
    _, a64, _, _ = oracle.domaths(os_, [payload], 64)
    e = relerr(a2, a64[0])
    say("  spectra %d err vs fp64 %.3e pcal framed vs stripped %.3e" % (m2, e, ep))
    # both take the integer-count fold (framed words mapped to their frames):
    # the same counts in the same order, so the phase-cal is bit-identical
    assert e <= TOL and np.array_equal(p2, p1[0])


def test_framed_host_run_counts_payload_spectra(hip):
    """Host-mode run on whole Mk5B frames whose frame bytes would hold one
    FFT more than their payload: 400 frames = 4,000,000 payload bytes = 625
    FFTs of 6400 4-channel time samples (5 spectra of 125), while 4,006,400
    frame bytes would give 626.  The core reports and copies back exactly the
    5 payload spectra (a guard after the host arrays stays untouched) and
    they equal the run on the stripped payload bit for bit."""
    from swspec_amd._lib import FLAG_MK5B_FRAMED
    n, fs, K = 6400, 32e6, 125
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * (K + 0.25) / fs, FFToverlapFactor=1,
                            WindowType="Hann", BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=4,
                            SourceFormat="Mk5B", ExtractPCal=True, MaxSourceBufferMB=4096, UseFile1Channel=2)
    assert s.core_overlapped_ffts == K and s.extract_pcal == 1
    payload = np.random.default_rng(5).integers(0, 256, 400 * 10000, dtype=np.uint8)
    frames = synth.mk5b_frames(payload)
    assert len(frames) // n == 626 and len(payload) // n == 625
    m, auto, _, pc, c1 = run_device(s, [payload])
    c1.finalize()
    assert m == 5
    core = hip.TaskCoreHIP(1)
    core.prepare(s, flags=FLAG_MK5B_FRAMED)
    inb = hip.Buffer(len(frames), pinned=True)
    inb.getData()[:] = frames
    inb.setLength(len(frames))
    ssb, T, guard = s.fft_ssb_points, s.pcal_tonebins, 4096
    out = hip.Buffer(m * 4 * ssb + guard)
    pcb = hip.Buffer(m * 8 * T + guard)
    for b in (out, pcb):
        b.getData()[:] = 0xAB
    core.run([inb], [out], [], [pcb])
    assert core.join() == m
    core.finalize()
    assert out.getLength() == m * 4 * ssb and pcb.getLength() == m * 8 * T
    assert (out.getData()[m * 4 * ssb:] == 0xAB).all() and (pcb.getData()[m * 8 * T:] == 0xAB).all()
    got = out.getData()[: m * 4 * ssb].view(np.float32).reshape(m, ssb)
    assert np.array_equal(got, auto[0])
    v = pcb.getData()[: m * 8 * T].view(np.float32)
    assert np.array_equal((v[0::2] + 1j * v[1::2]).reshape(m, T), pc[0].astype(np.complex64))


@pytest.mark.parametrize("bw,off,nch,fmt", [(64e6, 500e3, 1, "RawSigned"), (64e6, 500e3, 4, "Mk5B"),
                                            (128e6, 1e6, 1, "RawSigned")])
def test_pcal_fold_bins_beyond_thread_count(hip, oracle, bw, off, nch, fmt):
    """Phase-cal geometries whose integer-count fold runs fewer threads than
    tone bins: fs = 128 Ms/s with a 500 kHz offset (T = 128 bins, rotator
    L = 256, 64 threads) and fs = 256 Ms/s with a 1 MHz offset (T = L = 256,
    64 threads).  Every bin is written: the folded vector matches the FP64
    fold with the FP32 rotator (<= 1e-6) and the oracle's doMaths."""
    fs = 2 * bw
    n, nf = 25600, 3
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * (nf + 0.25) / fs, FFToverlapFactor=1,
                            WindowType="Hann", BandwidthHz=bw, BitsPerSample=2, SourceChannels=nch, SourceFormat=fmt,
                            ExtractPCal=True, PCalOffsetHz=off, MaxSourceBufferMB=4096, UseFile1Channel=nch)
    T, L = s.pcal_tonebins, s.pcal_rotatorlen
    say("pcal T=%d L=%d %s %d ch" % (T, L, fmt, nch))
    assert s.extract_pcal == 1 and T == int(fs / 1e6) and L == int(fs / np.gcd(int(fs), int(off)))
    rng = np.random.default_rng(int(off) % 997 + nch)
    comb = synth.pcal_comb(fs, bw, offset=off, amp=0.2, rng=rng)
    if nch == 1:
        raw = synth.raw_1ch(9, n * nf, tones=comb, fs=fs)
    else:
        raw = synth.raw_multi(9, nch, n * nf, tones_per_ch=[comb] * nch, fs=fs)
    m, _, _, pc, core = run_device(s, [raw])
    core.finalize()
    os_ = oracle_settings(oracle, s)
    _, _, _, p64 = oracle.domaths(os_, [raw], 64)
    x = oracle.unpack(os_, raw, n * nf, nch - 1).astype(np.float64)
    acc = np.zeros(2 * T)
    for f in range(nf):
        acc = oracle.pcal_fold64(x[f * n:(f + 1) * n], fs, off, T, L, rot32=True, acc=acc)
    ep = relerr(pc[0][0], acc[0::2] + 1j * acc[1::2])
    e64 = relerr(pc[0][0], p64[0][0])
    say("  pcal %.3e vs fp64 rotator %.3e" % (ep, e64))
    assert m == 1 and np.isfinite(pc[0][0]).all() and ep <= 1e-6 and e64 <= 1e-5


# ---------------------------------------------------------------- phase-cal at 3.2M with overlap

@pytest.mark.parametrize("fmt,nch,channels,framed", [("RawSigned", 1, [0], False), ("Mk5B", 4, [0, 1, 2, 3], False),
                                                    ("Mk5B", 8, [1, 6], False), ("Mk5B", 4, [2, 3], True),
                                                    ("Mk5B", 4, [1], False), ("Mk5B", 8, [4, 5, 6, 7], False)])
def test_pcal_3p2m_overlap_partial_vs_fold64(hip, oracle, fmt, nch, channels, framed):
    """3.2M points with the 16-tone comb through every integer-count fold
    instantiation: the per-channel fold (1-channel 2-bit; one channel of a
    4-channel Mk5B byte; 8-channel Mk5B channels 1 and 6, which sit in
    different bytes) and the grouped fold that reads a byte once for all its
    channels (4 channels of one byte, 4-channel and 8-channel Mk5B; 2 channels,
    framed Mk5B).  Each channel's folded vector against the FP64 fold with the FP32 rotator
    (<= 1e-6), its tones, the oracle's doMaths phase-cal and spectrum; two
    spectra (overlap 2, 3 FFTs each) and a partial third, phase-cal on every
    second FFT of the run (TaskCoreIPP.cpp:633-693)."""
    from swspec_amd._lib import FLAG_MK5B_FRAMED
    n, fs = 3200000, 32e6
    say("pcal 3.2M overlap", fmt, nch, channels, framed)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * 2.1 / fs, FFToverlapFactor=2,
                            WindowType="Cosine2", BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=nch,
                            SourceFormat=fmt, ExtractPCal=True, MaxSourceBufferMB=4096)
    K = s.core_overlapped_ffts
    assert s.extract_pcal == 1 and s.pcal_tonebins == 32 and K == 3
    nsamp = n // 2 * (2 * (K + 1) + 1) + n   # two spectra, then 2 FFTs of a third
    rng = np.random.default_rng(90 + nch)
    tones = [synth.pcal_comb(fs, fs / 2, amp=0.141, rng=rng) for _ in range(nch)]
    if nch == 1:
        raw = synth.raw_1ch(91, nsamp, tones=tones[0], fs=fs)
    else:
        raw = synth.raw_multi(91, nch, nsamp, tones_per_ch=tones, fs=fs)
    data = synth.mk5b_frames(raw) if framed else raw
    m, auto, _, pc, core = run_device(s, [data], channels=channels, flags=FLAG_MK5B_FRAMED if framed else 0)
    for j, ch in enumerate(channels):
        os_ = oracle_settings(oracle, hip.derive_settings(**dict(s.keys, UseFile1Channel=ch + 1)))
        m64, a64, _, p64 = oracle.domaths(os_, [raw], 64)
        assert m == m64 == 3
        x = oracle.unpack(os_, raw, nsamp, ch).astype(np.float64)
        for sp in range(m):
            nff = K if sp < 2 else 2
            acc = np.zeros(64)
            for k in range(nff):
                # the reference folds FFTs whose running count (across
                # spectra) is a multiple of the overlap factor
                if (sp * K + k) % 2:
                    continue
                f0 = (sp * (K + 1) + k) * n // 2
                acc = oracle.pcal_fold64(x[f0:f0 + n], fs, 10e3, 32, 3200, rot32=True, acc=acc)
            pref = acc[0::2] + 1j * acc[1::2]
            ep = relerr(pc[j][sp], pref)
            ep64 = relerr(pc[j][sp], p64[0][sp])
            tg = core.pcal_tones(j, sp)
            et = relerr(tg, oracle.pcal_tones64(acc, len(tg), 1e6, fs))
            e = relerr(auto[j][sp], a64[0][sp])
            say("  ch%d spectrum %d: auto %.3e pcal %.3e (vs fp64 rotator %.3e) tones %.3e" % (ch, sp, e, ep, ep64,
                                                                                              et))
            assert e <= TOL and ep <= 1e-6 and ep64 <= 1e-5 and et <= 1e-6
    core.finalize()
