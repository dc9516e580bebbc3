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


# ---------------------------------------------------------------- the other one-byte-per-sample modes
# VLBA1_2-*-4-2 (16 tracks, vlba_decode_2bit_16track_fanout2_decimation1 :5147),
# VLBA1_4-*-4-2 (32 tracks, _2bit_32track_fanout4 :5907), VLBA1_1/2/4-*-8-1
# (8/16/32 tracks, _1bit_8track_fanout1 :1362, _1bit_16track_fanout2 :2021,
# _1bit_32track_fanout4 :2891).  Payload = 2500 x ntrack bytes, chunk =
# 2 x (2520 - 12) x ntrack bytes, frame word = fanout bytes.

MODES = [(1, 2), (2, 2), (4, 2), (1, 1), (2, 1), (4, 1)]     # (fanout, nbit)


def _word_positions(mod, fanout, want, k=3):
    """Sample offsets of the first k frame words with modulate[word] == want."""
    return [int(m) * fanout for m in np.nonzero(mod == want)[0][:k]]


# (fanout, nbit, word bytes, {channel: levels of the word's `fanout` samples})
# derived by hand from initluts (mark5_format_vlba.c:90-165) and each
# decoder's output assignments:
#   16 tracks 2-bit: lut2bit2 entry j reads sign bit j + 2*(j//2) in {0,1,4,5}
#     and magnitude bit 2 above; ch0 = byte0 entries 0,1; ch1 = byte0 2,3;
#     ch2 = byte1 0,1; ch3 = byte1 2,3.  Bytes 0x05, 0xC0: ch0 sample 0 has
#     sign 1, mag 1 -> l = 3 -> +HiMag; sample 1 bits 1,3 clear -> -HiMag;
#     ch3 reads bits 4,6 then 5,7 of 0xC0 -> l = 2, 2 -> -1, -1.
#   32 tracks 2-bit: reorder32 keeps 0xAA55AA55, moves 0x55005500 down 7,
#     0x00AA00AA up 7; ch0..3 = reordered bytes 0, 2, 1, 3; sample u reads
#     bits 2u (sign), 2u+1 (mag).  Word 00 01 00 00 (x = 0x100): bit 8 is in
#     0x55005500 -> bit 1 -> reordered byte 0 = 0x02 -> ch0 sample 0 l = 2 -> -1.
#     Word 00 00 01 00 (x = 0x10000): bit 16 kept -> byte 2 = 0x01 -> ch1
#     sample 0 l = 1 -> +1.  Word 02 00 00 00: bit 1 in 0x00AA00AA -> bit 8 ->
#     byte 1 = 0x01 -> ch2 sample 0 -> +1.
#   1-bit: lut1bit = {+1, -1}[bit]; 8 tracks: ch c = bit c; 16 tracks: ch c =
#     byte c//4, bit 2(c%4)+u; 32 tracks: ch c = byte c%4, bit 2u + c//4.
H = float(HIMAG)
LITERALS = [
    (2, 2, [0x05, 0xC0], {0: [H, -H], 1: [-H, -H], 2: [-H, -H], 3: [-1.0, -1.0]}),
    (4, 2, [0x00, 0x01, 0x00, 0x00], {0: [-1.0, -H, -H, -H], 1: [-H] * 4, 2: [-H] * 4, 3: [-H] * 4}),
    (4, 2, [0x00, 0x00, 0x01, 0x00], {0: [-H] * 4, 1: [1.0, -H, -H, -H], 2: [-H] * 4, 3: [-H] * 4}),
    (4, 2, [0x02, 0x00, 0x00, 0x00], {0: [-H] * 4, 1: [-H] * 4, 2: [1.0, -H, -H, -H], 3: [-H] * 4}),
    (1, 1, [0x05], {0: [-1.0], 1: [1.0], 2: [-1.0], 3: [1.0], 7: [1.0]}),
    (2, 1, [0x06, 0x80], {0: [1.0, -1.0], 1: [-1.0, 1.0], 3: [1.0, 1.0], 7: [1.0, -1.0]}),
    (4, 1, [0x01, 0x00, 0x80, 0x00], {0: [-1.0, 1.0, 1.0, 1.0], 2: [1.0] * 4, 6: [1.0, 1.0, 1.0, -1.0]}),
]


@pytest.mark.parametrize("fanout,nbit,word,want", LITERALS)
def test_mode_level_literals(fanout, nbit, word, want):
    """One frame word repeated over a whole payload: unmodulated words give
    the hand-derived levels, modulated words their negation."""
    mod = O.vlba_modulate()
    P = 2500 * 8 * fanout
    raw = np.tile(np.array(word, dtype=np.uint8), P // fanout)
    for ch, lv in want.items():
        x = O.vlba_unpack(raw, P, ch, fanout=fanout, nbit=nbit)
        for s in _word_positions(mod, fanout, 0):
            assert list(x[s:s + fanout]) == [np.float32(v) for v in lv], (ch, s)
        for s in _word_positions(mod, fanout, 1):
            assert list(x[s:s + fanout]) == [-np.float32(v) for v in lv], (ch, s)


def _reference_luts():
    """initluts (mark5_format_vlba.c:90-165) tables, built as the reference builds them."""
    l2 = [1.0, -1.0]
    l4 = [-H, 1.0, -1.0, H]
    lut1 = np.zeros((256, 8), np.float32)
    lut21 = np.zeros((256, 4), np.float32)
    lut22 = np.zeros((256, 4), np.float32)
    for b in range(256):
        for i in range(8):
            lut1[b, i] = l2[(b >> i) & 1]
        for i in range(4):
            s = 2 * i
            lut21[b, i] = l4[((b >> s) & 1) + (((b >> (s + 1)) & 1) << 1)]
            s = i + (i // 2) * 2
            lut22[b, i] = l4[((b >> s) & 1) + (((b >> (s + 2)) & 1) << 1)]
    return lut1, lut21, lut22


def _table_decode(word, fanout, nbit, luts):
    """All channels of one frame word as the decoder's output assignments
    write them: returns [channel][sample]."""
    lut1, lut21, lut22 = luts
    if nbit == 2 and fanout == 1:
        fp = lut21[word[0]]
        return [[fp[c]] for c in range(4)]
    if nbit == 2 and fanout == 2:
        f0, f1 = lut22[word[0]], lut22[word[1]]
        return [[f0[0], f0[1]], [f0[2], f0[3]], [f1[0], f1[1]], [f1[2], f1[3]]]
    if nbit == 2 and fanout == 4:
        x = int(word[0]) | (int(word[1]) << 8) | (int(word[2]) << 16) | (int(word[3]) << 24)
        bits = (x & 0xAA55AA55) | ((x & 0x55005500) >> 7) | ((x & 0x00AA00AA) << 7)
        fp = [lut21[(bits >> (8 * k)) & 0xFF] for k in range(4)]
        return [[fp[0][u] for u in range(4)], [fp[2][u] for u in range(4)],
                [fp[1][u] for u in range(4)], [fp[3][u] for u in range(4)]]
    if fanout == 1:
        fp = lut1[word[0]]
        return [[fp[c]] for c in range(8)]
    if fanout == 2:
        f0, f1 = lut1[word[0]], lut1[word[1]]
        return [[f0[2 * c], f0[2 * c + 1]] for c in range(4)] + [[f1[2 * c], f1[2 * c + 1]] for c in range(4)]
    fp = [lut1[word[k]] for k in range(4)]
    return [[fp[c][2 * u] for u in range(4)] for c in range(4)] + \
           [[fp[c][2 * u + 1] for u in range(4)] for c in range(4)]


@pytest.mark.parametrize("fanout,nbit", MODES)
def test_modes_match_table_driven_decode(fanout, nbit):
    """Random words (no fill pattern) across two chunks: the oracle's bit
    arithmetic equals a decode through the reference's own LUT construction
    and output assignments, modulation restarting at every payload and
    chunk."""
    mod = O.vlba_modulate().astype(bool)
    luts = _reference_luts()
    T = 8 * fanout
    P, CH = 2500 * T, 2 * (2520 - 12) * T
    n = CH + 3 * P // 2
    rng = np.random.default_rng(40 + fanout * 3 + nbit)
    raw = _payload(rng, n + P)
    nch = 4 if nbit == 2 else 8
    want = np.zeros((nch, n), np.float32)
    for c0 in range(0, n, CH):
        for o in range(c0, min(c0 + CH, n), fanout):
            i = (o - c0) % P
            m = i // fanout
            vals = _table_decode(raw[o:o + fanout], fanout, nbit, luts)
            for c in range(nch):
                for u in range(fanout):
                    want[c, o + u] = -vals[c][u] if mod[m] else vals[c][u]
    for c in range(nch):
        assert np.array_equal(O.vlba_unpack(raw, n, c, avail=n + P, fanout=fanout, nbit=nbit), want[c]), c


def test_multizone_blank_hand_traces():
    """Payloads of 16 and 32 tracks hold 2 and 3 zones of 4096 words.
    16 tracks, fill from byte 36000 to the payload end (40000): zone 0 is all
    data; zone 1 (words 4096..4999, n = 904) bisects from i = 452 in steps
    226, 113, 56, 28, 14, 7, 3, 1 through words 4548 F, 4322 D, 4435 D, 4491 D,
    4519 F, 4505 F, 4498 D, 4501 F -> r = word 4501 = byte 36008; the 32-bit
    word at 36004 is fill -> end 36004.
    32 tracks, fill over bytes 32768..70000: zone 0 data, zone 1 (words
    4096..8191) all fill, zone 2 (words 8192..9999, n = 1808) findfirstvalid
    from i = 904: words 9096 D, 8644 F, 8870 D, 8757 D, 8701 F, 8729 F, 8743 F,
    8750 D, 8747 F -> r = word 8750 = byte 70000; the 32-bit word 8 bytes
    earlier (69992) is fill -> start 70004."""
    rng = np.random.default_rng(12)
    b = _payload(rng, 40000)
    _fill(b, 36000, 40000)
    assert O.vlba_blank_zone(b, 0, 40000, 0) == (0, 1 << 30)
    assert O.vlba_blank_zone(b, 0, 40000, 1) == (0, 36004)
    b = _payload(rng, 80000)
    _fill(b, 32768, 70000)
    assert O.vlba_blank_zone(b, 0, 80000, 0) == (0, 1 << 30)
    assert O.vlba_blank_zone(b, 0, 80000, 1) == (1 << 30, 0)
    assert O.vlba_blank_zone(b, 0, 80000, 2) == (70004, 1 << 30)
    # the decode: zone 1 of the 16-track payload zeroes words whose first
    # byte is at or past 36004, zone 2 of the 32-track one words before 70004
    b16 = _payload(rng, 80256 + 40000)
    _fill(b16, 36000, 40000)
    x = O.vlba_unpack(b16, 80256, 1, avail=len(b16), fanout=2, nbit=2)
    assert np.all(x[36004:40000] == 0) and np.all(x[:36004] != 0) and np.all(x[40000:80256] != 0)
    b32 = _payload(rng, 160512 + 80000)
    _fill(b32, 32768, 70000)
    x = O.vlba_unpack(b32, 160512, 5, avail=len(b32), fanout=4, nbit=1)
    assert np.all(x[32768:70004] == 0) and np.all(x[:32768] != 0) and np.all(x[70004:160512] != 0)


@pytest.mark.parametrize("fanout,nbit", MODES)
def test_factory_picks_vlba_modes(fanout, nbit):
    nch = 4 if nbit == 2 else 8
    s = O.derive(source_format="VLBA1_%d-256-%d-%d" % (fanout, nch, nbit), bits_per_sample=nbit,
                 source_channels=nch, fft_points=40000, fft_integ_seconds=0.01, bandwidth_hz=8e6,
                 fft_overlap_factor=1)
    assert s.vlba_fanout == fanout
    assert O.lib().so_pick_unpacker(O.C.byref(s)) == 4
    assert s.raw_fullfft_bytes == 40000
