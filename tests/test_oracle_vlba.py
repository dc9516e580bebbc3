"""Pins of the oracle's VLBA decoder (VLBAUnpacker over mark5access,
src/IA-32/DataUnpackers.cpp:459-535; mark5access/mark5access/
mark5_format_vlba.c:90-196, 4537-4583; blanker_mark5.c:41-200).

mark5access cannot be built here (its sources include an autotools-generated
config.h), so no decoded output of the reference exists to compare with.
These tests pin the restatement by what the mathematics and the source's
constants fix: the modulation recurrence and its period, literal level
tables, blank zones of synthetic fill patterns found by linear scan, and the
chunk/payload restart of the modulation index."""
import numpy as np
import pytest

from oracle import oracle as O

HIMAG = np.float32(3.3359)
FILL = bytes([0x44, 0x33, 0x22, 0x11])   # 0x11223344 little-endian


def _recurrence(n):
    """k(t) = k(t-11) ^ k(t-13) ^ k(t-14) ^ k(t-16), k(-1..-16) = 1: the
    register of initmodulate written as a linear recurrence."""
    k = [1] * 16 + [0] * n
    for t in range(16, 16 + n):
        k[t] = k[t - 11] ^ k[t - 13] ^ k[t - 14] ^ k[t - 16]
    return np.array(k[16:], dtype=np.uint8)


def test_modulation_is_the_recurrence_with_every_ninth_step_skipped():
    mod = O.vlba_modulate()
    raw = _recurrence(20000 + 20000 // 8 + 8)
    # outputs i use raw step i + floor(i/8): one step per 8 outputs is the parity bit
    idx = np.arange(20000) + np.arange(20000) // 8
    assert np.array_equal(mod, raw[idx])


def test_modulation_register_is_maximal_length():
    # x^16 + x^14 + x^13 + x^11 + 1 is primitive: period 65535, balanced to one
    seq = _recurrence(2 * 65535)
    assert np.array_equal(seq[:65535], seq[65535:])
    assert int(seq[:65535].sum()) == 32768
    for p in (3, 5, 17, 257):   # proper divisors of 65535 = 3*5*17*257
        assert not np.array_equal(seq[: 65535 // p], seq[65535 // p: 2 * (65535 // p)])


def _unmodulated_positions(mod, want):
    return np.nonzero(mod == want)[0][:4]


@pytest.mark.parametrize("byte,levels", [
    (0x00, [-HIMAG, -HIMAG, -HIMAG, -HIMAG]),
    (0xFF, [HIMAG, HIMAG, HIMAG, HIMAG]),
    (0x1B, [HIMAG, -1.0, 1.0, -HIMAG]),     # 00 01 10 11: ch0 = bits 0,1 = 11
    (0xE4, [-HIMAG, 1.0, -1.0, HIMAG]),     # 11 10 01 00
    (0x55, [1.0, 1.0, 1.0, 1.0]),           # sign bit set, magnitude clear
    (0xAA, [-1.0, -1.0, -1.0, -1.0]),
])
def test_level_table_literals(byte, levels):
    """lut4level = {-HiMag, 1, -1, HiMag}[bit(2c) + 2 bit(2c+1)], negated where
    the modulation bit is 1 (Q2: byte 0x00 is -3.3359 in mark5access)."""
    mod = O.vlba_modulate()
    raw = np.full(20000, byte, dtype=np.uint8)
    for ch in range(4):
        x = O.vlba_unpack(raw, 20000, ch)
        for i in _unmodulated_positions(mod, 0):
            assert x[i] == np.float32(levels[ch])
        for i in _unmodulated_positions(mod, 1):
            assert x[i] == -np.float32(levels[ch])


def _payload(rng, n=20000):
    b = rng.integers(0, 256, n, dtype=np.uint8)
    # keep random data free of accidental fill words
    b[::4] = np.where(b[::4] == 0x44, 0x45, b[::4])
    return b


def _fill(buf, a, b):
    for o in range(a, b, 4):
        buf[o:o + 4] = np.frombuffer(FILL, dtype=np.uint8)


def _scan_blank(buf, seg):
    """Linear scan: index of the first fill word / first data word (8-byte words)."""
    words = buf[seg:seg + 20000].reshape(-1, 8)
    isfill = np.all(words == np.frombuffer(FILL * 2, dtype=np.uint8), axis=1)
    return isfill


def test_blank_none_and_all():
    rng = np.random.default_rng(1)
    b = _payload(rng)
    assert O.vlba_blank(b, 0) == (0, 1 << 30)
    _fill(b, 0, 20000)
    assert O.vlba_blank(b, 0) == (1 << 30, 0)


@pytest.mark.parametrize("word", [8, 625, 1250, 2000, 2499])
def test_blank_tail_fill_found_by_scan(word):
    """Data then fill from `word`: where the bisection of findfirstinvalid
    (steps 625, 312, ..., 1 from word 1250) lands on the boundary, the data
    zone ends at the first fill word a linear scan finds."""
    rng = np.random.default_rng(word)
    b = _payload(rng)
    _fill(b, 8 * word, 20000)
    isfill = _scan_blank(b, 0)
    first = int(np.argmax(isfill)) * 8
    assert first == 8 * word
    assert O.vlba_blank(b, 0) == (0, first)


def test_blank_bisection_known_answers():
    """Hand traces of blanker_mark5.c's bisections where they miss the
    boundary (the restatement must miss it the same way).
    Tail fill from word 7: probes 1250, 625, 313, 157, 79, 40, 21, 12, 8 are
    fill, 6 is data -> r = word 8 = byte 64; the 32-bit word at byte 60 is
    fill -> 60.  Head fill on words 0..5: findfirstvalid ends at r = word 6 =
    byte 48; the 32-bit word at byte r - 8 = 40 is fill -> 52."""
    rng = np.random.default_rng(9)
    b = _payload(rng)
    _fill(b, 56, 20000)
    assert O.vlba_blank(b, 0) == (0, 60)
    b = _payload(rng)
    _fill(b, 0, 48)
    assert O.vlba_blank(b, 0) == (52, 1 << 30)


def test_blank_head_fill_keeps_offset_of_straggler_rule():
    """Fill then data from word w: findfirstvalid's straggler test looks at
    the 32-bit word 8 bytes before the boundary, which is always fill, so the
    data zone starts 4 bytes into the first data word when the bisection is
    exact (w = 1250, 625: probed directly)."""
    for w in (1250, 625):
        rng = np.random.default_rng(w + 1)
        b = _payload(rng)
        _fill(b, 0, 8 * w)
        assert O.vlba_blank(b, 0) == (8 * w + 4, 1 << 30)


def test_blank_32bit_straggler():
    """Fill starting 4 bytes into a word: the data zone ends at that 4-byte
    boundary (findfirstinvalid's data32[r/4 - 1] check)."""
    rng = np.random.default_rng(3)
    b = _payload(rng)
    _fill(b, 8 * 1249 + 4, 20000)      # first whole fill word 1250: the bisection's first probe
    assert O.vlba_blank(b, 0) == (0, 8 * 1249 + 4)


def test_blanked_samples_are_zero_and_others_decode():
    rng = np.random.default_rng(4)
    b = _payload(rng, 40128)
    _fill(b, 12000, 20000)        # tail of the first payload
    x = O.vlba_unpack(b, 40128, 2)
    assert np.all(x[12000:20000] == 0)
    assert np.all(x[:12000] != 0) and np.all(x[20000:] != 0)


def test_modulation_restarts_every_payload_and_chunk():
    """A chunk is 40128 bytes = payloads at 0, 20000 and a 128-byte tail at
    40000; each starts at modulation index 0 (Q7)."""
    rng = np.random.default_rng(5)
    n = 3 * 40128 + 1000
    b = _payload(rng, n + 20000)
    mod = O.vlba_modulate().astype(bool)
    x = O.vlba_unpack(b, n, 1)
    lvl = np.array([-HIMAG, 1.0, -1.0, HIMAG], dtype=np.float32)
    base = lvl[(b[:n] >> 2) & 3]
    o = np.arange(n)
    i = (o % 40128) % 20000
    want = np.where(mod[i], -base, base)
    assert np.array_equal(x, want)


def test_factory_picks_vlba_for_4ch_2bit():
    s = O.derive(source_format="VLBA", bits_per_sample=2, source_channels=4, fft_points=40000,
                 fft_integ_seconds=0.01, bandwidth_hz=8e6, fft_overlap_factor=1)
    assert O.lib().so_pick_unpacker(O.C.byref(s)) == 4
    assert s.raw_fullfft_bytes == 40000


def test_domaths_vlba_matches_unpack_then_rfft():
    """doMaths on VLBA bytes: the spectrum of the decoded samples (Hann,
    one FFT per spectrum, 1/K = 1)."""
    n = 50000
    s = O.derive(source_format="VLBA", bits_per_sample=2, source_channels=4, fft_points=n,
                 fft_integ_seconds=n / 16e6, bandwidth_hz=8e6, fft_overlap_factor=1, window="Hann",
                 use_channel_file1=3, max_source_buffer_mb=64)
    rng = np.random.default_rng(6)
    raw = _payload(rng, 2 * n)
    _fill(raw, 30000, 40000)                 # a blank zone inside the first FFT's second payload
    nsp, auto, _, _ = O.domaths(s, [raw], precision=64)
    x = O.vlba_unpack(raw, n, 2).astype(np.float64) * O.window("Hann", n, fp64=True)
    want = np.abs(np.fft.rfft(x)) ** 2
    got = np.asarray(auto[0]).reshape(-1, n // 2 + 1)[0]
    assert nsp >= 1
    assert np.max(np.abs(got - want)) <= 1e-9 * np.max(want)
