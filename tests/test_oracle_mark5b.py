"""Pins of the oracle's Mark5B decoder (Mark5BUnpacker over mark5access,
src/IA-32/DataUnpackers.cpp:660-758; mark5access/mark5access/
mark5_format_mark5b.c:59-100 (initluts), :239-1911 (decimation-1 decoders);
blanker_mark5.c:126-200).

mark5access cannot be built here (its sources include an autotools-generated
config.h).  The restatement is pinned by literal bytes worked by hand from
initluts and the decoders' output assignments, by a decode through tables
built as initluts builds them, by blank zones per 10000-byte payload, and by
the fixed pattern that stands in for the reference's std::rand() samples."""
import numpy as np
import pytest

from oracle import oracle as O

H = float(np.float32(3.3359))
FILL = np.frombuffer(bytes([0x44, 0x33, 0x22, 0x11]), dtype=np.uint8)
# (nchan, nbit) that Mark5BUnpacker::canHandleConfig admits and mark5access decodes
CONFIGS = [(2, 1), (4, 1), (8, 1), (16, 1), (32, 1), (2, 2), (4, 2), (8, 2), (16, 2)]


def _data(rng, n):
    b = rng.integers(0, 256, n, dtype=np.uint8)
    b[::4] = np.where(b[::4] == 0x44, 0x45, b[::4])
    return b


# (nchan, nbit, repeated group bytes, {channel: levels of the group's samples})
# worked by hand: lut2bit entry i = {-HiMag, 1, -1, HiMag}[bit 2i + 2 bit 2i+1],
# lut1bit entry i = {+1, -1}[bit i]; 4 channels x 2 bits = one sample per
# byte, entries 0..3 = channels 0..3; 2 channels x 1 bit = four samples per
# byte, data[c][o] = fp[2u + c]; 16 channels x 2 bits = 4 bytes per sample,
# channel c from byte c/4, entry c%4.
LITERALS = [
    (4, 2, [0x1B], {0: [H], 1: [-1.0], 2: [1.0], 3: [-H]}),
    (2, 1, [0x06], {0: [1.0, -1.0, 1.0, 1.0], 1: [-1.0, 1.0, 1.0, 1.0]}),
    (16, 2, [0x00, 0x00, 0x03, 0x00], {8: [H], 9: [-H], 0: [-H], 15: [-H]}),
    (8, 1, [0x81], {0: [-1.0], 1: [1.0], 7: [-1.0]}),
    (2, 2, [0x36], {0: [-1.0, H], 1: [1.0, -H]}),   # entries -1, +1, +H, -H; ch0 = fp[0], fp[2]
]


@pytest.mark.parametrize("nchan,nbit,group,want", LITERALS)
def test_level_literals(nchan, nbit, group, want):
    B = nchan * nbit
    spg = max(8 // B, 1)                        # samples per group
    count = 64 * spg
    raw = np.tile(np.array(group, dtype=np.uint8), 64)
    for ch, lv in want.items():
        x = O.mark5b_unpack(raw, count, ch, nchan, nbit)
        for g in (16, 17, 40):                  # groups past the 16 substituted samples
            assert list(x[g * spg:(g + 1) * spg]) == [np.float32(v) for v in lv], (ch, g)


def test_first_16_samples_are_the_fixed_pattern():
    """The reference writes std::rand() signs x 3.3359 into samples 0..15 of
    every request (DataUnpackers.cpp:747-750); the restatement writes
    +3.3359, -3.3359, ... whatever the bytes."""
    raw = np.zeros(400, dtype=np.uint8)
    for nchan, nbit in CONFIGS:
        x = O.mark5b_unpack(raw, 40, 0, nchan, nbit)
        assert list(x[:16]) == [np.float32(H if n % 2 == 0 else -H) for n in range(16)]


def _reference_luts():
    l2 = [1.0, -1.0]
    l4 = [-H, 1.0, -1.0, H]
    lut1 = np.array([[l2[(b >> i) & 1] for i in range(8)] for b in range(256)], np.float32)
    lut2 = np.array([[l4[((b >> (2 * i)) & 1) + (((b >> (2 * i + 1)) & 1) << 1)] for i in range(4)]
                     for b in range(256)], np.float32)
    return lut1, lut2


@pytest.mark.parametrize("nchan,nbit", CONFIGS)
def test_decode_matches_table_driven_decode(nchan, nbit):
    """Random bytes over three payloads: the oracle's bit arithmetic equals a
    decode through initluts' tables in the decoders' order (B < 8: one byte
    per iteration, 8/B samples, entry u*nchan + c; B >= 8: B/8 bytes per
    sample, channel c from byte c/(8/nbit), entry c % (8/nbit))."""
    lut1, lut2 = _reference_luts()
    lut = lut1 if nbit == 1 else lut2
    per = 8 // nbit                              # channels per byte
    B = nchan * nbit
    count = (3 * 10000 * 8) // B
    raw = _data(np.random.default_rng(B + nbit), count * B // 8 + 16)
    want = np.zeros((nchan, count), np.float32)
    if B < 8:
        spb = 8 // B
        for i in range(count // spb):
            fp = lut[raw[i]]
            for u in range(spb):
                for c in range(nchan):
                    want[c, i * spb + u] = fp[u * nchan + c]
    else:
        g = B // 8
        for t in range(count):
            for c in range(nchan):
                want[c, t] = lut[raw[t * g + c // per]][c % per]
    want[:, :16] = [H if n % 2 == 0 else -H for n in range(16)]
    for c in range(nchan):
        assert np.array_equal(O.mark5b_unpack(raw, count, c, nchan, nbit), want[c]), c


def test_blank_zone_per_payload():
    """Each 10000-byte payload gets its own blank zone (one 1250-word zone):
    a fill over the end of payload 1 blanks the samples whose byte group
    starts there; payload 2 restarts with data.  4 channels x 2 bits: sample
    t = byte t."""
    raw = _data(np.random.default_rng(3), 40000)
    for o in range(18000, 20000, 4):
        raw[o:o + 4] = FILL
    x = O.mark5b_unpack(raw, 30000, 2, 4, 2)
    assert np.all(x[18000:20000] == 0)
    assert np.all(x[16:18000] != 0) and np.all(x[20000:30000] != 0)
    # 16 channels x 2 bits: 4 bytes per sample, the same bytes -> samples 4500..4999
    y = O.mark5b_unpack(raw, 7500, 5, 16, 2)
    assert np.all(y[4500:5000] == 0) and np.all(y[16:4500] != 0) and np.all(y[5000:] != 0)


def test_factory_and_domaths():
    """Mark5B-<Mbps>-4-2 is picked as decoder 5; doMaths over it is the
    spectrum of the decoded samples."""
    n = 30000
    s = O.derive(source_format="Mark5B-512-4-2", bits_per_sample=2, source_channels=4, fft_points=n,
                 fft_integ_seconds=n / 16e6, bandwidth_hz=8e6, fft_overlap_factor=1, window="Hann",
                 use_channel_file1=2, max_source_buffer_mb=64)
    assert O.lib().so_pick_unpacker(O.C.byref(s)) == 5
    assert s.raw_fullfft_bytes == n
    raw = _data(np.random.default_rng(8), 2 * n)
    nsp, auto, _, _ = O.domaths(s, [raw], precision=64)
    x = O.mark5b_unpack(raw, n, 1, 4, 2, avail=len(raw)).astype(np.float64) * O.window("Hann", n, fp64=True)
    want = np.abs(np.fft.rfft(x)) ** 2
    assert nsp >= 1
    assert np.max(np.abs(auto[0][0] - want)) <= 1e-9 * np.max(want)
