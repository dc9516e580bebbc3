"""Pins of the CPU oracle against known answers that do not come from the
oracle's own formulas: hand-computed configuration tables, literal decoder
tables, numpy/pocketfft, Parseval, the reference's phase-cal unit-test KAT
(src/IA-32/PhaseCal/PCal.cpp:861-951) and the design test vector of
src/matlab/swspectrometer.m (cosine window, 50% overlap, tone of period 37).
"""
import numpy as np
import pytest

from swspec_amd import synth

M5 = np.float32(3.3359)


# ---------------------------------------------------------------- settings

def test_settings_32000pt_1ch(oracle):
    s = oracle.derive(fft_points=32000, fft_integ_seconds=1.6, fft_overlap_factor=1, window="Hann",
                      bandwidth_hz=16e6, bits_per_sample=2, source_channels=1)
    assert (s.averaged_ffts, s.raw_fullfft_bytes, s.raw_overlap_bytes) == (1600, 8000, 8000)
    assert (s.rawbuf_size, s.max_spectra_per_buffer, s.max_buffers_per_spectrum) == (25600000, 2, 0)
    assert s.core_overlapped_ffts == 1600 and s.fft_ssb_points == 16001


@pytest.mark.parametrize("cap,rawbuf,S,B,K", [(-1.0, 12800000, 4, 25, 4), (1024.0, 320000000, 100, 0, 100)])
def test_settings_3p2m_4ch_mk5b(oracle, cap, rawbuf, S, B, K):
    s = oracle.derive(fft_points=3200000, fft_integ_seconds=10.0, fft_overlap_factor=1, window="Cosine2",
                      bandwidth_hz=16e6, bits_per_sample=2, source_channels=4, source_format="Mk5B",
                      extract_pcal=1, max_source_buffer_mb=cap)
    assert s.averaged_ffts == 100 and s.raw_fullfft_bytes == 3200000
    assert (s.rawbuf_size, s.max_specffts_per_buffer, s.max_buffers_per_spectrum) == (rawbuf, S, B)
    assert s.core_overlapped_ffts == K
    assert (s.extract_pcal, s.pcal_tonebins, s.pcal_rotatorlen, s.pcal_pulses) == (1, 32, 3200, 100)


@pytest.mark.parametrize("n,T", [(1 << 20, 327.68), (1000000, 312.5)])
def test_settings_xpol_pair(oracle, n, T):
    s = oracle.derive(fft_points=n, fft_integ_seconds=T, fft_overlap_factor=1, bandwidth_hz=32e6,
                      bits_per_sample=2, source_channels=1, num_sources=2, do_xpol=1)
    assert s.averaged_ffts == 20000 and s.num_xpols == 1
    if n == 1 << 20:
        assert (s.rawbuf_size, s.max_buffers_per_spectrum, s.core_overlapped_ffts) == (8388608, 625, 32)


def test_settings_defaults_overlap(oracle):
    # reference defaults: 320000 points, 20 s, 50% overlap, 16 MHz, 8-bit
    s = oracle.derive()
    assert (s.averaged_ffts, s.raw_overlap_bytes, s.raw_fullfft_bytes) == (1000, 160000, 320000)
    assert (s.rawbuf_size, s.max_specffts_per_buffer, s.max_buffers_per_spectrum) == (2560000, 8, 125)
    assert s.core_overlapped_ffts == 15     # 2*8 - 1


def test_settings_float_arithmetic(oracle):
    # 32e6 * 0.7f rounds to 22400000 in float -> 700 FFTs; double would give 699
    s = oracle.derive(fft_points=32000, fft_integ_seconds=0.7, fft_overlap_factor=1, bandwidth_hz=16e6,
                      bits_per_sample=2)
    assert s.averaged_ffts == 700
    assert int(32e6 * np.float64(np.float32(0.7)) / 32000) == 699


def test_settings_pcal_disable(oracle):
    # PCal needs N % rotatorlen == 0 (swspectrometer.cpp:185-189): 4096 % 3200 != 0
    s = oracle.derive(fft_points=4096, fft_integ_seconds=1.0, fft_overlap_factor=1, bandwidth_hz=16e6,
                      bits_per_sample=2, extract_pcal=1)
    assert s.extract_pcal == 0
    s = oracle.derive(fft_points=320000, fft_integ_seconds=1.0, fft_overlap_factor=1, bandwidth_hz=32e6,
                      bits_per_sample=2, extract_pcal=1)
    assert (s.extract_pcal, s.pcal_tonebins, s.pcal_rotatorlen) == (1, 64, 6400)


def test_settings_errors(oracle):
    with pytest.raises(ValueError):
        oracle.derive(use_channel_file1=2, source_channels=1)
    with pytest.raises(ValueError):
        oracle.derive(extract_pcal=1, pcal_offset_hz=0.0)
    assert oracle.gcd(32000000, 10000) == 10000 and oracle.gcd(64000000, 10000) == 10000


def test_product_derive_matches_oracle(oracle):
    """swspec_hip_derive (product C++) against the oracle restatement."""
    import swspec_amd
    cases = [
        dict(FFTpoints=32000, FFTIntegrationTimeSec=1.6, FFToverlapFactor=1, BandwidthHz=16e6, BitsPerSample=2),
        dict(FFTpoints=3200000, FFTIntegrationTimeSec=10, FFToverlapFactor=1, BandwidthHz=16e6, BitsPerSample=2,
             SourceChannels=4, SourceFormat="Mk5B", ExtractPCal=True),
        dict(FFTpoints=3200000, FFTIntegrationTimeSec=10, FFToverlapFactor=1, BandwidthHz=16e6, BitsPerSample=2,
             SourceChannels=4, SourceFormat="Mk5B", ExtractPCal=True, MaxSourceBufferMB=1024),
        dict(),
        dict(FFTpoints=32000, FFTIntegrationTimeSec=0.7, FFToverlapFactor=1, BandwidthHz=16e6, BitsPerSample=2),
        dict(FFTpoints=1 << 20, FFTIntegrationTimeSec=327.68, FFToverlapFactor=1, BandwidthHz=32e6,
             BitsPerSample=2, NumCPUCores=3),
    ]
    omap = dict(FFTpoints="fft_points", FFTIntegrationTimeSec="fft_integ_seconds",
                FFToverlapFactor="fft_overlap_factor", BandwidthHz="bandwidth_hz", BitsPerSample="bits_per_sample",
                SourceChannels="source_channels", SourceFormat="source_format", ExtractPCal="extract_pcal",
                MaxSourceBufferMB="max_source_buffer_mb", NumCPUCores="num_cores")
    for kw in cases:
        p = swspec_amd.derive_settings(**kw)
        okw = {omap[k]: (int(v) if isinstance(v, bool) else v) for k, v in kw.items()}
        o = oracle.derive(**okw)
        for f in ["averaged_ffts", "raw_fullfft_bytes", "raw_overlap_bytes", "rawbuf_size",
                  "max_spectra_per_buffer", "max_specffts_per_buffer", "max_buffers_per_spectrum",
                  "core_overlapped_ffts", "extract_pcal", "pcal_tonebins", "pcal_rotatorlen"]:
            assert getattr(p, f) == getattr(o, f), (kw, f)


# ---------------------------------------------------------------- decoders

def _s(oracle, fmt="RawSigned", nch=1, incr=1):
    return oracle.derive(bits_per_sample=2, source_channels=nch, source_format=fmt, channelorder_increasing=incr,
                         fft_points=32000, fft_overlap_factor=1, fft_integ_seconds=1.0)


def test_unpack_1ch_literals(oracle):
    s = _s(oracle)
    raw = np.array([0x00, 0xFF, 0x1B, 0xE4] * 4, dtype=np.uint8)
    x = oracle.unpack(s, raw, 64, 0)
    exp = np.array([[1, 1, 1, 1], [-M5, -M5, -M5, -M5], [1, -1, M5, -M5], [-M5, M5, -1, 1]] * 4,
                   dtype=np.float32).reshape(-1)
    assert np.array_equal(x[:64], exp)


def test_unpack_4ch_literals(oracle):
    raw = np.array([0xE4] * 8 + [0x1B] * 8, dtype=np.uint8)
    inc = _s(oracle, nch=4, incr=1)
    dec = _s(oracle, nch=4, incr=0)
    # increasing: channel c in bits 2c..2c+1 -> 0xE4 = 11 10 01 00
    for c, v in enumerate([1, -1, M5, -M5]):
        assert np.all(oracle.unpack(inc, raw, 16, c)[:8] == np.float32(v))
    # decreasing: channel c in bits 6-2c..7-2c
    for c, v in enumerate([-M5, M5, -1, 1]):
        assert np.all(oracle.unpack(dec, raw, 16, c)[:8] == np.float32(v))
    mk = _s(oracle, fmt="Mk5B", nch=4, incr=0)     # Mk5BUnpacker ignores the order key
    for c, v in enumerate([1, -1, M5, -M5]):
        assert np.all(oracle.unpack(mk, raw, 16, c)[:8] == np.float32(v))


def test_unpack_8ch_byte_select(oracle):
    # 8 channels: 2 bytes per time sample, channel c in byte c/4
    raw = np.array([0x00, 0xFF] * 8, dtype=np.uint8)
    inc = _s(oracle, nch=8, incr=1)
    dec = _s(oracle, nch=8, incr=0)
    for c in range(8):
        assert np.all(oracle.unpack(inc, raw, 8, c) == np.float32(1.0 if c < 4 else -M5))
        assert np.all(oracle.unpack(dec, raw, 8, c) == np.float32(-M5 if c < 4 else 1.0))


def test_unpack_matches_synth_quantiser(oracle):
    """Synthetic packers + decoder round trip: decode(pack(quantize(x))) = level(quantize(x))."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((4, 4096))
    f = np.stack([synth.quantize(v) for v in x])
    s = _s(oracle, nch=4)
    raw = synth.pack_multi(f)
    for c in range(4):
        assert np.array_equal(oracle.unpack(s, raw, 4096, c), synth.LEVELS[f[c]])
    raw1 = synth.pack_1ch(f[0])
    assert np.array_equal(oracle.unpack(_s(oracle), raw1, 4096, 0), synth.LEVELS[f[0]])
    # quantiser thresholds at +-0.98 sigma
    q = synth.quantize(np.array([0.98, 0.979, 0.0, -1e-9, -0.98, -0.981]))
    assert list(q) == [2, 0, 0, 1, 1, 3]


def test_unpack_unsupported(oracle):
    s = oracle.derive(bits_per_sample=2, source_channels=2, source_format="Mk5B", fft_points=32000)
    assert oracle.lib().so_pick_unpacker(s) == 0


# ---------------------------------------------------------------- windows

def test_windows_closed_forms(oracle):
    for n in [16, 1024, 32000]:
        i = np.arange(n)
        assert np.array_equal(oracle.window("None", n), np.ones(n, dtype=np.float32))
        # Cosine2 = sin^2(pi (2i+1) / (2N)) (same function, different expression)
        c2 = oracle.window("Cosine2", n, fp64=True)
        assert np.allclose(c2, np.sin(np.pi * (2 * i + 1) / (2 * n)) ** 2, rtol=0, atol=1e-15)
        c1 = oracle.window("Cosine", n, fp64=True)
        assert np.allclose(c1, np.sin(np.pi * (2 * i + 1) / (2 * n)), rtol=0, atol=1e-15)
        assert np.allclose(oracle.window("Hann", n, fp64=True), np.hanning(n), atol=1e-15)
        assert np.allclose(oracle.window("Hamming", n, fp64=True), np.hamming(n), atol=1e-15)
        # IPP Blackman(alpha=0.5) - numpy Blackman(alpha=-0.16) = 0.66 sin^2(2 pi n/(N-1))
        d = oracle.window("Blackman", n, fp64=True) - np.blackman(n)
        assert np.allclose(d, 0.66 * np.sin(2 * np.pi * i / (n - 1)) ** 2, atol=1e-14)
        # FP32 table is the FP64 value rounded once
        assert np.array_equal(oracle.window("Cosine2", n), c2.astype(np.float32))


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [8, 12, 30, 42, 64, 100, 1000, 4096, 20000, 32000, 65536, 320000, 1 << 20, 1000000])
def test_rfft64_vs_numpy(oracle, n):
    x = np.random.default_rng(n).standard_normal(n)
    X = oracle.rfft64(x)
    R = np.fft.rfft(x)
    assert np.abs(X - R).max() <= 1e-13 * np.abs(R).max()
    assert X[0].imag == 0 and X[-1].imag == 0
    # Parseval for a real sequence
    p = (np.abs(X[0]) ** 2 + np.abs(X[-1]) ** 2 + 2 * np.sum(np.abs(X[1:-1]) ** 2)) / n
    assert abs(p - np.sum(x * x)) <= 1e-10 * np.sum(x * x)


def test_rfft64_brute_force(oracle):
    for n in [2, 4, 6, 10, 14, 18, 30, 48]:
        x = np.random.default_rng(n).standard_normal(n)
        k = np.arange(n // 2 + 1)[:, None]
        D = np.exp(-2j * np.pi * k * np.arange(n)[None, :] / n) @ x
        assert np.abs(oracle.rfft64(x) - D).max() < 1e-12


def test_rfft32_accuracy(oracle):
    for n in [32000, 3200000]:
        x = np.random.default_rng(1).standard_normal(n).astype(np.float32)
        e = np.linalg.norm(oracle.rfft32(x) - np.fft.rfft(x.astype(np.float64))) / np.linalg.norm(np.fft.rfft(x))
        assert e < 5e-7


def test_cfft64_vs_numpy(oracle):
    z = np.random.default_rng(3).standard_normal(3000) + 1j * np.random.default_rng(4).standard_normal(3000)
    assert np.abs(oracle.cfft64(z) - np.fft.fft(z)).max() < 1e-10


# ---------------------------------------------------------------- doMaths

def _one_channel(oracle, n=32000, T=None, nffts=8, **kw):
    fs = 32e6
    T = T if T is not None else n * nffts / fs
    args = dict(fft_points=n, fft_integ_seconds=T, fft_overlap_factor=1, window="Hann", bandwidth_hz=fs / 2,
                bits_per_sample=2, source_channels=1)
    args.update(kw)
    return oracle.derive(**args)


def test_domaths_tone_bin(oracle):
    """32000-point tone at 0.3*fs/2 = 4.8 MHz lands on bin 4800 (df = 1 kHz)."""
    s = _one_channel(oracle, nffts=4)
    raw = synth.raw_1ch(11, 32000 * 4, tones=[(0.3, 4.8e6, 0.0)], fs=32e6)
    n, auto, _, _ = oracle.domaths(s, [raw], 64)
    assert n == 1
    assert int(np.argmax(auto[0][0][1:-1])) + 1 == 4800


def test_domaths_design_vector(oracle):
    """src/matlab/swspectrometer.m: 1024 points, cosine window, 50% overlap,
    noise + 0.1 cos(2 pi n / 37): peak at bin round(1024/37) = 28."""
    n = 1024
    s = oracle.derive(fft_points=n, fft_integ_seconds=n * 256 / 2e6, fft_overlap_factor=2, window="Cosine",
                      bandwidth_hz=1e6, bits_per_sample=2, source_channels=1, max_source_buffer_mb=1024)
    K = s.core_overlapped_ffts
    assert K == 2 * 256 - 1
    rng = np.random.default_rng(37)
    x = rng.standard_normal(n * 256) + 0.1 * np.cos(2 * np.pi * np.arange(n * 256) / 37)
    raw = synth.pack_1ch(synth.quantize(x))
    m, auto, _, _ = oracle.domaths(s, [raw], 64)
    assert m == 1
    assert int(np.argmax(auto[0][0][1:-1])) + 1 == 28


def test_domaths_fp32_vs_fp64(oracle):
    s = _one_channel(oracle, nffts=16)
    raw = synth.raw_1ch(2, 32000 * 16, tones=[(0.05, 4.8e6, 0.0)], fs=32e6)
    _, a32, _, _ = oracle.domaths(s, [raw], 32)
    _, a64, _, _ = oracle.domaths(s, [raw], 64)
    e = np.linalg.norm(a32[0] - a64[0]) / np.linalg.norm(a64[0])
    assert e < 1e-6


def test_domaths_matches_numpy_pipeline(oracle):
    """FP64 doMaths = mean over FFTs of |rfft(window * unpack)|^2 with numpy."""
    n, nf = 4096, 6
    s = oracle.derive(fft_points=n, fft_integ_seconds=n * nf / 16e6, fft_overlap_factor=1, window="Cosine2",
                      bandwidth_hz=8e6, bits_per_sample=2, source_channels=4, source_format="Mk5B",
                      max_source_buffer_mb=64)
    raw = synth.raw_multi(9, 4, n * nf, fs=16e6)
    _, auto, _, _ = oracle.domaths(s, [raw], 64)
    x = synth.LEVELS[(raw >> 4) & 3].astype(np.float64)   # channel 0? no: UseFile1Channel=1 -> channel 0
    x = synth.LEVELS[raw & 3].astype(np.float64)
    w = np.sin(np.pi * (2 * np.arange(n) + 1) / (2 * n)) ** 2
    P = np.mean([np.abs(np.fft.rfft(x[i * n:(i + 1) * n] * w)) ** 2 for i in range(nf)], axis=0)
    assert np.abs(auto[0][0] - P).max() <= 1e-12 * P.max()


def test_domaths_xpol_equal_inputs(oracle):
    s = _one_channel(oracle, n=4000, nffts=3, num_sources=2, do_xpol=1)
    raw = synth.raw_1ch(4, 4000 * 3, fs=32e6)
    n, auto, xpol, _ = oracle.domaths(s, [raw, raw], 64)
    assert n == 1
    assert np.allclose(xpol[0][0].real, auto[0][0], rtol=1e-12, atol=0)
    assert np.abs(xpol[0][0].imag).max() <= 1e-12 * auto[0][0].max()
    assert np.array_equal(auto[0], auto[1])


def test_domaths_combine_identity(oracle):
    """Sum_b (1/K) Sum_{k in b} |X_k|^2 = (1/K) Sum_all (TaskCoreIPP.cpp:124,311-349)."""
    n = 4000
    big = _one_channel(oracle, n=n, nffts=4, max_source_buffer_mb=1024)
    # cap so that two FFTs fit a buffer: K = 2, two buffers per spectrum
    small = _one_channel(oracle, n=n, nffts=4, max_source_buffer_mb=2 * (n // 4) * 2 / 1048576 * 1.0 + 1e-9)
    assert big.core_overlapped_ffts == 4
    raw = synth.raw_1ch(8, n * 4, fs=32e6)
    _, a, _, _ = oracle.domaths(big, [raw], 64)
    bb = n // 4 * small.core_overlapped_ffts
    comb = sum(oracle.domaths(small, [raw[i:i + bb]], 64)[1][0][0] for i in range(0, n, bb))
    K = small.core_overlapped_ffts
    assert np.allclose(comb, a[0][0] * 4 / K, rtol=1e-12)


def test_domaths_partial_and_overlap(oracle):
    n = 4000
    s = _one_channel(oracle, n=n, nffts=3, max_source_buffer_mb=1024)
    raw = synth.raw_1ch(6, n * 4, fs=32e6)
    m, auto, _, _ = oracle.domaths(s, [raw], 64)
    assert m == 2          # one full spectrum of 3 FFTs + a partial of 1, unscaled
    x = synth.LEVELS[np.stack([(raw >> 6) & 3, (raw >> 4) & 3, (raw >> 2) & 3, raw & 3], 1).reshape(-1)]
    w = np.hanning(n)
    P3 = np.abs(np.fft.rfft(x[3 * n:4 * n] * w)) ** 2
    assert np.allclose(auto[0][1], P3, rtol=1e-10, atol=1e-9)
    so = _one_channel(oracle, n=n, nffts=3, fft_overlap_factor=2, max_source_buffer_mb=1024)
    assert so.core_overlapped_ffts == 5
    assert oracle.count_spectra(so, n // 4 * 3) == (1, 5)


# ---------------------------------------------------------------- phase-cal

def _kat_data(samplecount, bw, spacing, offset, sampleoffset, ntones):
    n = np.arange(samplecount, dtype=np.float64)
    x = np.zeros(samplecount)
    for t in range(ntones):
        x += np.sin(np.pi * (n + sampleoffset) * (offset + t * spacing) / bw + t * np.pi * 5 / 180)
    return x


def test_pcal_kat(oracle):
    """PCal.cpp:861-951 with './test 32000 16e6 1e6 510e3 0': phases -90 + 5 t deg within 0.1 deg,
    including the skip-3-samples continuation via the sample offset."""
    samplecount, bw, sp, off, so = 32000, 16e6, 1e6, 510e3, 0
    fs = 2 * bw
    nt = oracle.pcal_ntones(bw, sp, off)
    assert nt == 16
    nbins = int(fs / sp)
    L = int(fs / oracle.gcd(int(fs), int(off)))
    x = _kat_data(samplecount, bw, sp, off, so, nt)
    acc = oracle.pcal_fold64(x, fs, off, nbins, L, sampleoffset=so)
    acc = oracle.pcal_fold64(x[3:], fs, off, nbins, L, sampleoffset=so + samplecount + 3, acc=acc)
    tones = oracle.pcal_tones64(acc, nt, sp, fs)
    ph = np.degrees(np.angle(tones))
    assert np.all(np.abs(ph - (-90 + 5 * np.arange(nt))) < 0.1)
    ref = oracle.pcal_reference64(x, bw, sp, off, so)
    assert np.all(np.abs(np.degrees(np.angle(ref)) - (-90 + 5 * np.arange(nt))) < 0.1)


def test_pcal_core_fold(oracle):
    """extract_PCal FP32 restatement vs FP64 fold with the FP32 rotator."""
    s = oracle.derive(fft_points=320000, fft_integ_seconds=0.01, fft_overlap_factor=1, bandwidth_hz=16e6,
                      bits_per_sample=2, extract_pcal=1)
    x = synth.LEVELS[synth.quantize(np.random.default_rng(2).standard_normal(320000))].astype(np.float32)
    p32 = np.zeros(2 * s.pcal_tonebins, dtype=np.float32)
    oracle.lib().so_extract_pcal32(s, x.ctypes.data, 320000, p32.ctypes.data)
    p64 = oracle.pcal_fold64(x.astype(np.float64), 32e6, 10e3, 32, 3200, rot32=True)
    assert np.abs(p32 - p64).max() < 1e-3 * np.abs(p64).max() + 1e-2
    # pulses of one bin: sum over n = b mod 32 of x[n] rot[n mod 3200]
    b = 5
    idx = np.arange(b, 320000, 32)
    rot = np.exp(-2j * np.pi * 10e3 * (idx % 3200) / 32e6)
    direct = np.sum(x[idx].astype(np.float64) * rot)
    assert abs(complex(p64[2 * b], p64[2 * b + 1]) - direct) < 1e-6 * abs(direct) + 1e-6


# ---------------------------------------------------------------- Mk5B framing

def test_mk5b_strip(oracle):
    payload = np.random.default_rng(1).integers(0, 256, 30000, dtype=np.uint8)
    frames = synth.mk5b_frames(payload)
    assert len(frames) == 3 * 10016
    assert np.array_equal(oracle.mk5b_strip(frames), payload)
    bad = frames.copy()
    bad[10016] ^= 1
    with pytest.raises(ValueError):
        oracle.mk5b_strip(bad)
