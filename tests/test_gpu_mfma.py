"""GPU parity of the radix-64 split-FP16 MFMA first column stage
(SWSPEC_FLAG_MFMA64) against the FP64 oracle and the plain FP32 path.

The reference has no such stage: IPP's DFT (TaskCoreIPP.cpp:427-434) runs in
FP32, so the bar is the FP32 path's own: normwise relative error <= 1e-5 of
every output (auto spectra, cross spectrum) vs the FP64 oracle on the same
seeded input.
"""
import numpy as np
import pytest

from gpu_util import oracle_settings, relerr, run_device, say
from swspec_amd import synth
from swspec_amd._lib import FLAG_MFMA64

pytestmark = pytest.mark.gpu
TOL = 1e-5

# (N, window, FFTs): column lengths 512 (2^20), 1280 with a partly filled
# column tile (3.2M), 128 with one column per workgroup (32000), 64 where the
# MFMA stage is the whole column FFT (8192, 4096), 320 (320000); 640 (819200)
# and 1024 (2^21) with a composite radix (10, 16) after the MFMA stage
SIZES = [(1 << 20, "Cosine", 2), (3200000, "Cosine2", 2), (65536, "Hann", 3), (32000, "Hann", 4),
         (8192, "Hamming", 5), (4096, "Blackman", 5), (320000, "Blackman", 2),
         (819200, "Hann", 2), (1 << 21, "Cosine2", 2)]


@pytest.mark.parametrize("n,win,nf", SIZES)
def test_mfma64_spectrum_vs_oracle(hip, oracle, n, win, nf):
    say("mfma64 spectrum", n, win, nf)
    info = hip.plan_info(n, FLAG_MFMA64)
    assert info["mfma64"] == 1 and info["p1_radix0"] == 64
    fs = 32e6
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType=win,
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, MaxSourceBufferMB=4096)
    raw = synth.raw_1ch(n + nf + 5, n * nf, tones=[(0.2, 4.8e6, 0.3)], fs=fs)
    m, auto, _, _, core = run_device(s, [raw], flags=FLAG_MFMA64)
    core.finalize()
    m0, auto0, _, _, core0 = run_device(s, [raw])
    core0.finalize()
    os_ = oracle_settings(oracle, s)
    m64, a64, _, _ = oracle.domaths(os_, [raw], 64)
    assert m == m0 == m64 == 1
    e64, e0 = relerr(auto[0], a64[0]), relerr(auto0[0], a64[0])
    say("  N1 %d B %d: mfma vs fp64 %.3e (fp32 path %.3e), mfma vs fp32 path %.3e" % (
        info["N1"], info["B"], e64, e0, relerr(auto[0], auto0[0])))
    assert e64 <= TOL and relerr(auto[0], auto0[0]) <= TOL


@pytest.mark.parametrize("n,nf", [(1 << 20, 3), (32000, 12), (3200000, 2)])
def test_mfma64_noise_floor_accuracy(hip, oracle, n, nf):
    """Noise-only spectra (no tone to dominate the norm): the MFMA path must
    match the FP32 path's accuracy, <= 1e-6 vs FP64.  Regression for split
    inputs on an FP16 rounding tie, whose hi and lo once came from two
    different roundings (one FP16 ulp off: 3e-6 to 7e-6 here)."""
    say("mfma64 noise floor", n, nf)
    fs = 64e6
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, MaxSourceBufferMB=4096)
    raw = synth.pack_1ch(synth.quantize(np.random.default_rng(nf).standard_normal(n * nf)))
    _, a64, _, _ = oracle.domaths(oracle_settings(oracle, s), [raw], 64)
    m, auto, _, _, core = run_device(s, [raw], flags=FLAG_MFMA64)
    core.finalize()
    e = relerr(auto[0], a64[0])
    say("  mfma vs fp64 %.3e" % e)
    assert m == 1 and e <= 1e-6


def _hilbert(x):
    """Imaginary part of the analytic signal (FP64 FFT)."""
    n = len(x)
    h = np.zeros(n)
    h[0] = 1.0
    h[1:n // 2] = 2.0
    h[n // 2] = 1.0
    return np.fft.ifft(np.fft.fft(x) * h).imag


def test_mfma64_xpol_2p20_pair(hip, oracle):
    """Cross-pol pair at the 2^20-point length: X, Y Gaussian with complex
    correlation 0.2 at 30 degrees (Y = rho (cos X + sin H{X}) + sqrt(1 - rho^2) Z),
    2-bit, 64 Ms/s, Kahan FP32 accumulators; error of XX, YY and XY* vs FP64
    with the MFMA stage on and off."""
    n, nf, fs, rho = 1 << 20, 3, 64e6, 0.2
    say("mfma64 xpol", n, nf)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType="Hann",
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, DoCrossPolarization=True,
                            MaxSourceBufferMB=4096, num_sources=2)
    rng = np.random.default_rng(20)
    a = rng.standard_normal(n * nf)
    ph = np.deg2rad(30.0)
    b = rho * (np.cos(ph) * a + np.sin(ph) * _hilbert(a)) + np.sqrt(1 - rho ** 2) * rng.standard_normal(n * nf)
    ra, rb = synth.pack_1ch(synth.quantize(a)), synth.pack_1ch(synth.quantize(b))
    os_ = oracle_settings(oracle, s)
    _, a64, x64, _ = oracle.domaths(os_, [ra, rb], 64)
    errs = {}
    for name, fl in (("mfma", FLAG_MFMA64), ("fp32", 0)):
        m, auto, xp, _, core = run_device(s, [ra, rb], flags=fl)
        core.finalize()
        assert m == 1
        errs[name] = (relerr(auto[0], a64[0]), relerr(auto[1], a64[1]), relerr(xp[0], x64[0]))
        assert xp[0][0, 0].imag == 0 and xp[0][0, -1].imag == 0
    for name, (ex, ey, exy) in errs.items():
        say("  %s: XX %.3e YY %.3e XY* %.3e" % (name, ex, ey, exy))
    assert max(errs["mfma"]) <= TOL and max(errs["fp32"]) <= TOL
    # the 30-degree correlation phase survives 2-bit quantisation in the band average
    xy = x64[0][0, 1:-1].sum()
    say("  XY* phase %.2f deg" % np.degrees(np.angle(xy)))


def test_mfma64_multichannel_pcal_matches_fp32_path(hip, oracle):
    """4-channel Mk5B with phase-cal at the 3.2M-point length: the MFMA stage
    changes only the spectra (within 1e-5 of the FP64 oracle), never the
    phase-cal vector (bit-identical: it does not pass through the FFT)."""
    n, nf, fs = 3200000, 2, 32e6
    say("mfma64 mk5b pcal", n, nf)
    s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1,
                            WindowType="Cosine2", BandwidthHz=16e6, BitsPerSample=2, SourceChannels=4,
                            SourceFormat="Mk5B", ExtractPCal=True, PCalOffsetHz=10000.0, MaxSourceBufferMB=4096)
    raw = np.random.default_rng(5).integers(0, 256, n * nf, dtype=np.uint8)
    m, auto, _, pc, core = run_device(s, [raw], channels=[1], flags=FLAG_MFMA64)
    core.finalize()
    m0, auto0, _, pc0, core0 = run_device(s, [raw], channels=[1])
    core0.finalize()
    os_ = oracle_settings(oracle, hip.derive_settings(**dict(s.keys, UseFile1Channel=2)))
    _, a64, _, _ = oracle.domaths(os_, [raw], 64)
    e64 = relerr(auto[0], a64[0])
    say("  spectrum mfma vs fp64 %.3e, fp32 path %.3e" % (e64, relerr(auto0[0], a64[0])))
    assert m == m0 == 1 and e64 <= TOL
    assert np.array_equal(pc[0], pc0[0])


def test_mfma64_rejects_lengths_without_64_columns(hip):
    """20000 = 2 x 2^4 x 5^4 has no column length that is a multiple of 64."""
    from swspec_amd import TaskCoreHIP
    fs = 32e6
    s = hip.derive_settings(FFTpoints=20000, FFTIntegrationTimeSec=20000 * 2 / fs, FFToverlapFactor=1,
                            BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, MaxSourceBufferMB=4096)
    core = TaskCoreHIP(0)
    with pytest.raises(RuntimeError):
        core.prepare(s, flags=FLAG_MFMA64)
