"""GPU: short FFTs (M = N/2 <= 8192 complex points) run in one pass, k_short:
decode + window + the whole M-point FFT in LDS + r2c split + detection, with
each accumulator slice's sums in registers.  Same parity bar as the two-pass
path (normwise relative error <= 1e-5 against the FP64 oracle restating
TaskCoreIPP::doMaths, src/IA-32/TaskCoreIPP.cpp:394-524), over the edge
lengths of the plan (N = 16, odd M, the 16/32 points-per-thread choice, the
8192-point bound), long chunks, multi-channel Mk5B with phase-cal, cross-power
pairs (up to 128 KB of LDS per workgroup), framed input and the mark5access
decodes."""
import numpy as np
import pytest

from gpu_util import oracle_settings, relerr, run_device, say
from swspec_amd import synth
from test_gpu_core import _vlba_bytes

pytestmark = pytest.mark.gpu

TOL = 1e-5
FS = 32e6


def _single(hip, n):
    info = hip.plan_info(n)
    assert info["single"] == 1 and info["N1"] == n // 2 and info["N2"] == 1, info
    return info


# (N, window, FFTs, format): N = 16 (M = 8, one thread per FFT), N = 30
# (M = 15 odd, radices 5, 3), N = 2430 (M = 1215 odd), N = 2048 (M = 1024,
# 32 points per thread), N = 4096 (M = 2048, 16 points per thread), N = 12800
# (M = 6400), N = 16384 (M = 8192: the bound, 32 points per thread).  Odd M
# needs a format with whole bytes per FFT: 4-channel Mk5B (channel 3).
SHORT = [(16, "Hann", 7, "raw1"), (30, "None", 5, "mk5b4"), (2430, "Cosine2", 6, "mk5b4"),
         (2048, "Blackman", 9, "raw1"), (4096, "Cosine2", 40, "raw1"), (12800, "Cosine", 5, "mk5b4"),
         (16384, "Hamming", 6, "raw1")]


@pytest.mark.parametrize("n,win,nf,fmt", SHORT)
def test_short_spectrum_vs_oracle(hip, oracle, n, win, nf, fmt):
    info = _single(hip, n)
    say("short", n, win, nf, fmt, "points/thread", info["p1_maxpts"], "threads", info["p1_threads"])
    mk = fmt == "mk5b4"
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / FS, FFToverlapFactor=1, WindowType=win,
                            BandwidthHz=FS / 2, BitsPerSample=2, SourceChannels=4 if mk else 1,
                            SourceFormat="Mk5B" if mk else "RawSigned", UseFile1Channel=4 if mk else 1,
                            MaxSourceBufferMB=4096)
    assert s.core_overlapped_ffts == nf
    tone = [(0.2, 4.8e6, 0.3)]
    raw = (synth.raw_multi(n + nf, 4, n * nf, tones_per_ch=[tone] * 4, fs=FS) if mk
           else synth.raw_1ch(n + nf, n * nf, tones=tone, fs=FS))
    m, auto, _, _, core = run_device(s, [raw])
    core.finalize()
    os_ = oracle_settings(oracle, s)
    m64, a64, _, _ = oracle.domaths(os_, [raw], 64)
    _, a32, _, _ = oracle.domaths(os_, [raw], 32)
    assert m == m64 == 1
    e64, e32 = relerr(auto[0], a64[0]), relerr(auto[0], a32[0])
    say("  err vs fp64 %.3e  vs fp32 restatement %.3e" % (e64, e32))
    assert e64 <= TOL and e32 <= TOL


def test_short_long_chunk_overlap_and_partial(hip, oracle):
    """4096-point FFTs, overlap 2, two full spectra of 1499 overlapped FFTs
    (0.096 s at 32 Ms/s) and a partial one: each spectrum is one chunk spread
    over 374 accumulator slices of 4 FFTs (and the partial is written
    unscaled, TaskCoreIPP.cpp:519-524)."""
    n = 4096
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=0.096, FFToverlapFactor=2,
                            WindowType="Hann", BandwidthHz=FS / 2, BitsPerSample=2, SourceChannels=1,
                            MaxSourceBufferMB=4096)
    K = s.core_overlapped_ffts
    assert K == 1499
    # a spectrum walks (K + 1) half-FFTs; then 500 FFTs of a third
    nsamp = n // 2 * (2 * (K + 1) + 500) + n
    raw = synth.raw_1ch(77, nsamp - nsamp % 4, tones=[(0.1, 1.3e6, 0.0)], fs=FS)
    m, auto, _, _, core = run_device(s, [raw])
    core.finalize()
    os_ = oracle_settings(oracle, s)
    m64, a64, _, _ = oracle.domaths(os_, [raw], 64)
    assert m == m64 == 3
    errs = [relerr(auto[0][i], a64[0][i]) for i in range(m)]
    say("  spectra %d errs %s" % (m, ["%.2e" % e for e in errs]))
    assert max(errs) <= TOL


def test_short_mk5b_channels_and_pcal(hip, oracle):
    """6400-point FFTs (M = 3200) of 4-channel Mk5B with the phase-cal comb:
    all four channels in one launch, spectra and phase-cal against the
    oracle."""
    n, nf = 6400, 30
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / FS, FFToverlapFactor=1,
                            WindowType="Cosine2", BandwidthHz=FS / 2, BitsPerSample=2, SourceChannels=4,
                            SourceFormat="Mk5B", ExtractPCal=True, MaxSourceBufferMB=4096)
    _single(hip, n)
    rng = np.random.default_rng(64)
    tones = [synth.pcal_comb(FS, FS / 2, amp=0.141, rng=rng) for _ in range(4)]
    raw = synth.raw_multi(9, 4, n * nf, tones_per_ch=tones, fs=FS)
    m, auto, _, pc, core = run_device(s, [raw], channels=[0, 1, 2, 3])
    core.finalize()
    assert m == 1
    for ch in range(4):
        os_ = oracle_settings(oracle, hip.derive_settings(**dict(s.keys, UseFile1Channel=ch + 1)))
        _, a64, _, p64 = oracle.domaths(os_, [raw], 64)
        e, ep = relerr(auto[ch], a64[0]), relerr(pc[ch], p64[0])
        say("  ch%d spectrum %.3e pcal %.3e" % (ch, e, ep))
        assert e <= TOL and ep <= 1e-5


@pytest.mark.parametrize("n", [8192, 16384])
def test_short_xpol_vs_oracle(hip, oracle, n):
    """Cross-power pairs in one pass: both sources' FFTs of a slice in one
    workgroup (16384 points: 2 x 64 KB of LDS)."""
    nf, fs = 12, 64e6
    _single(hip, n)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, DoCrossPolarization=True,
                            MaxSourceBufferMB=4096, num_sources=2)
    assert s.num_xpols == 1
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n * nf)
    b = 0.3 * a + np.sqrt(1 - 0.09) * rng.standard_normal(n * nf)
    ra, rb = synth.pack_1ch(synth.quantize(a)), synth.pack_1ch(synth.quantize(b))
    m, auto, xp, _, core = run_device(s, [ra, rb])
    core.finalize()
    os_ = oracle_settings(oracle, s)
    _, a64, x64, _ = oracle.domaths(os_, [ra, rb], 64)
    assert m == 1
    ea, eb, ex = relerr(auto[0], a64[0]), relerr(auto[1], a64[1]), relerr(xp[0], x64[0])
    say("  n %d auto %.3e %.3e xpol %.3e" % (n, ea, eb, ex))
    assert ea <= TOL and eb <= TOL and ex <= TOL
    assert xp[0][0, 0].imag == 0 and xp[0][0, -1].imag == 0


def test_short_framed_equals_stripped(hip):
    """Whole Mk5B frames decoded by the single pass: spectra and phase-cal
    bit-identical to the run on the stripped payload."""
    from swspec_amd._lib import FLAG_MK5B_FRAMED
    n, nf = 3200, 50
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / FS, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=FS / 2, BitsPerSample=2, SourceChannels=4, SourceFormat="Mk5B",
                            ExtractPCal=True, MaxSourceBufferMB=4096, UseFile1Channel=3)
    payload = np.random.default_rng(12).integers(0, 256, 16 * 10000, dtype=np.uint8)
    m1, a1, _, p1, c1 = run_device(s, [payload])
    c1.finalize()
    m2, a2, _, p2, c2 = run_device(s, [synth.mk5b_frames(payload)], flags=FLAG_MK5B_FRAMED)
    c2.finalize()
    assert m1 == m2 == 1
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(p1[0], p2[0])


@pytest.mark.parametrize("fmt,nch,nbit,n", [("VLBA1_2-512-4-2", 4, 2, 8000), ("Mark5B-512-8-2", 8, 2, 4000)])
def test_short_mark5access_vs_oracle(hip, oracle, fmt, nch, nbit, n):
    """mark5access decodes (blank-zone tables per FFT) through the single
    pass, overlapping FFTs, fill zones inside FFTs."""
    nf = 12
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / FS, FFToverlapFactor=2, WindowType="Hann",
                            BandwidthHz=FS / 2, BitsPerSample=nbit, SourceChannels=nch, SourceFormat=fmt,
                            MaxSourceBufferMB=4096, UseFile1Channel=nch - 1)
    _single(hip, n)
    nbytes = s.raw_fullfft_bytes * (nf + 3)
    raw = _vlba_bytes(np.random.default_rng(n + nch), nbytes,
                      [(1000, 5000), (nbytes // 3, nbytes // 3 + 9000), (nbytes - 20000, nbytes - 12000)])
    m, auto, _, _, core = run_device(s, [raw])
    core.finalize()
    os_ = oracle_settings(oracle, s)
    m64, a64, _, _ = oracle.domaths(os_, [raw], 64)
    assert m == m64
    e = relerr(auto[0], a64[0])
    say("  %s spectra %d err %.3e" % (fmt, m, e))
    assert e <= TOL
