"""Pins of the oracle's PCal extractor classes (src/IA-32/PhaseCal/PCal.cpp):
the factory's choices, the reference unit test (PCal.cpp:861-951), the
brute-force rotating fold of extractAndIntegrate_reference (:139-175), impulse
responses of the Shifting sample map and delayed synthetic combs for the
time-domain delay."""
import numpy as np
import pytest

from oracle import oracle as O


def unit_test_data(samplecount, bw, spacing, offset, sampleoffset, ntones):
    """PCal.cpp:898-908: sum of tones at offset + t*spacing with phase t*5 deg."""
    n = np.arange(samplecount, dtype=np.float64)
    d = np.zeros(samplecount)
    for t in range(ntones):
        d += np.sin(np.pi * (n + sampleoffset) * (offset + t * spacing) / bw + t * np.pi * 5 / 180)
    return d.astype(np.float32)


def run_unit_test(ex, samplecount, bw, spacing, offset, sampleoffset):
    """PCal.cpp:911-918: one extract, then the same data shifted by 3 samples
    after adjustSampleOffset(sampleoffset + samplecount + 3)."""
    d = unit_test_data(samplecount, bw, spacing, offset, sampleoffset, ex.getLength())
    assert ex.extractAndIntegrate(d)
    ex.adjustSampleOffset(sampleoffset + samplecount + 3)
    assert ex.extractAndIntegrate(d[3:])
    return ex.getFinalPCal()


def phase_err_deg(tones, first=0):
    ph = np.degrees(np.angle(tones))
    exp = -90.0 + 5.0 * np.arange(len(tones))
    return np.abs((ph - exp + 180) % 360 - 180)[first:]


@pytest.mark.parametrize("bw,sp,off,kind,nbins,ntones,L", [
    (16e6, 1e6, 510e3, "ImplicitShift", 3200, 16, 0),
    (16e6, 1e6, 10e3, "ImplicitShift", 3200, 16, 0),
    (32e6, 1e6, 10e3, "ImplicitShift", 6400, 32, 0),
    (8e6, 1e6, 10e3, "ImplicitShift", 1600, 8, 0),
    (16e6, 1e6, 0, "Trivial", 32, 16, 0),
    (16e6, 1e6, 256e3, "Shifting", 32, 16, 125),
    (16e6, 3e6, 10e3, "ImplicitShift", 3200, 6, 0),
])
def test_factory_choice(bw, sp, off, kind, nbins, ntones, L):
    ex = O.PCalExtractor(bw, sp, off)
    assert (ex.kind, ex.getNBins(), ex.getLength(), ex.s.rotatorlen) == (kind, nbins, ntones, L)


def test_reference_unit_test_implicit():
    # './test 32000 16e6 1e6 510e3 0' (the example of PCal.cpp:872)
    ex = O.PCalExtractor(16e6, 1e6, 510e3)
    assert ex.kind == "ImplicitShift"
    n, tones = run_unit_test(ex, 32000, 16e6, 1e6, 510e3, 0)
    assert n == 32000 + 31997
    assert phase_err_deg(tones).max() < 0.1
    assert np.allclose(np.abs(tones), 10.0, rtol=1e-3)   # 1/2 amplitude x 64000 samples / 3200 bins
    assert not ex.extractAndIntegrate(np.zeros(8, np.float32))   # finalized
    ex.clear()
    assert ex.extractAndIntegrate(np.zeros(8, np.float32))


def test_reference_unit_test_trivial():
    # offset 0: tone 0 sits at DC where the test signal is sin(0) = 0, so
    # only tones 1.. carry a phase
    ex = O.PCalExtractor(16e6, 1e6, 0)
    assert ex.kind == "Trivial"
    _, tones = run_unit_test(ex, 32000, 16e6, 1e6, 0, 0)
    assert abs(tones[0]) < 1e-3 * abs(tones[1])
    assert phase_err_deg(tones, first=1).max() < 0.1


def test_reference_unit_test_sampleoffset():
    ex = O.PCalExtractor(32e6, 1e6, 10e3, sampleoffset=777)
    _, tones = run_unit_test(ex, 64000, 32e6, 1e6, 10e3, 777)
    assert phase_err_deg(tones).max() < 0.1


@pytest.mark.parametrize("bw,sp,off,so", [(16e6, 1e6, 510e3, 0), (16e6, 1e6, 10e3, 5), (32e6, 1e6, 10e3, 11)])
def test_matches_bruteforce_rotating_fold(bw, sp, off, so):
    """ImplicitShift tones equal extractAndIntegrate_reference's (per-sample
    rotator, 2*ntones bins) scaled by nbins / (2*ntones)."""
    rng = np.random.default_rng(5)
    x = rng.choice(np.array([-3.3359, -1, 1, 3.3359], np.float32), 20011)
    ex = O.PCalExtractor(bw, sp, off, sampleoffset=so)
    for a, b in [(0, 7001), (7001, 7002), (7002, 20011)]:   # several calls, odd lengths
        ex.extractAndIntegrate(x[a:b])
    _, tones = ex.getFinalPCal()
    ref = O.pcal_reference64(x.astype(np.float64), bw, sp, off, sampleoffset=so)
    scale = ex.getNBins() / (2 * ex.getLength())
    assert np.abs(tones * scale - ref).max() <= 1e-9 * np.abs(ref).max()


def test_trivial_is_dft_of_samples():
    """Trivial tone k = (1/nbins) sum_n x[n] e^{-2 pi i k (n + so) / nbins}."""
    rng = np.random.default_rng(6)
    x = rng.standard_normal(5003).astype(np.float32)
    ex = O.PCalExtractor(16e6, 1e6, 0, sampleoffset=11)
    ex.extractAndIntegrate(x[:999])
    ex.extractAndIntegrate(x[999:])
    _, tones = ex.getFinalPCal()
    n = np.arange(len(x)) + 11
    ref = np.array([(x * np.exp(-2j * np.pi * k * n / 32)).sum() / 32 for k in range(16)])
    assert np.abs(tones - ref).max() <= 1e-11 * np.abs(ref).max()


def test_shifting_whole_rotator_matches_fold():
    """Shifting built directly with rotatorlen a multiple of nbins keeps
    every sample: equal to the generic FP32-rotator fold (so_pcal_fold64)."""
    bw, sp, off = 16e6, 1e6, 10e3
    ex = O.PCalExtractor(bw, sp, off, sampleoffset=3, kind="Shifting")
    assert ex.s.rotatorlen == 3200 and ex.getNBins() == 32
    rng = np.random.default_rng(9)
    x = rng.standard_normal(10000).astype(np.float32)
    ex.extractAndIntegrate(x[:4321])
    ex.extractAndIntegrate(x[4321:])
    _, tones = ex.getFinalPCal()
    acc = O.pcal_fold64(x.astype(np.float64), 2 * bw, off, 32, 3200, sampleoffset=3, rot32=True)
    ref = O.pcal_tones64(acc[0::2] + 1j * acc[1::2], 16, sp, 2 * bw)
    assert np.abs(tones - ref).max() <= 1e-12 * np.abs(ref).max()


def _rot32(off, fs, j):
    a = 2 * np.pi * (-off / fs) * j
    return complex(np.float32(np.cos(a)), np.float32(np.sin(a)))


@pytest.mark.parametrize("pos,kept", [(0, True), (40, True), (95, True), (96, False), (124, False),
                                      (125 + 17, True), (125 + 100, False), (250 + 5, True), (250 + 9, True)])
def test_shifting_drop_rule_impulses(pos, kept):
    """Factory Shifting at 256 kHz offset: rotatorlen 125, nbins 32.  In whole
    rotator periods only 3 pulses of 32 samples are added (PCal.cpp:432-441),
    the trailing partial period is added whole (:444-460)."""
    bw, sp, off = 16e6, 1e6, 256e3
    x = np.zeros(260, np.float32)   # 2 whole periods + a tail of 10
    x[pos] = 1.0
    ex = O.PCalExtractor(bw, sp, off)
    assert ex.kind == "Shifting"
    ex.extractAndIntegrate(x)
    _, tones = ex.getFinalPCal()
    if not kept:
        assert np.all(tones == 0)
        return
    j = pos % 125
    b = j % 32
    exp = _rot32(off, 2 * bw, j) / 32 * np.exp(-2j * np.pi * np.arange(16) * b / 32)
    assert np.allclose(tones, exp, rtol=0, atol=1e-12)


def test_implicit_tone_base_int_wrap():
    """nbins * offset is an int product (PCal.cpp:669): 6400 * 510000 wraps,
    the base bin is out of range and no tone is copied."""
    ex = O.PCalExtractor(32e6, 1e6, 510e3)
    assert ex.kind == "ImplicitShift" and ex.getNBins() == 6400
    ex.extractAndIntegrate(np.ones(6400, np.float32))
    seed = np.full(ex.getLength(), 7 + 7j)
    _, tones = ex.getFinalPCal(seed)
    assert np.all(tones == 7 + 7j)


@pytest.mark.parametrize("tau_frac", [0.0, 0.137, 0.5, 0.9031])
def test_delay_of_delayed_comb(tau_frac):
    sp, f0, nt = 1e6, 10e3, 16
    tau = tau_frac / sp
    f = f0 + sp * np.arange(nt)
    tones = 3.0 * np.exp(-2j * np.pi * f * tau + 0.4j)
    d = O.pcal_delay64(tones, sp)
    err = (d - tau + 0.5 / sp) % (1 / sp) - 0.5 / sp   # circular
    assert abs(err) < 1e-6 / sp
    assert 0 <= d < 1 / sp


def test_delay_from_extracted_comb():
    """A pulse train of period 32 samples starting at sample 37 at 32 Ms/s
    (1 MHz comb): the delay within the 1 us period is 37 mod 32 samples."""
    bw, sp, off = 16e6, 1e6, 0
    x = np.zeros(3200, np.float32)
    x[37::32] = 1.0
    ex = O.PCalExtractor(bw, sp, off)
    ex.extractAndIntegrate(x)
    _, tones = ex.getFinalPCal()
    d = O.pcal_delay64(tones[1:], sp)   # tones 1..15: the DC tone carries no delay
    assert abs(d - 5 / 32e6) < 1e-3 / sp
