"""Seeded synthetic baseband in the reference's on-disk layouts.

2-bit quantiser at +-0.98 sigma with the reference's field convention
(src/IA-32/DataUnpackers.cpp:212: field 0 -> +1, 1 -> -1, 2 -> +3.3359,
3 -> -3.3359):  x >= t -> 2,  0 <= x < t -> 0,  -t <= x < 0 -> 1,  x < -t -> 3.
Packings:
  raw 1-channel : 4 samples per byte, oldest sample in the MSB
                  (TwoBitSinglechannelUnpacker, DataUnpackers.cpp:267-268)
  raw/Mk5B 4n-ch: channel c in bits 2*(c%4) of byte c/4 of each time sample
                  (TwoBitUnpacker increasing order, Mk5BUnpacker, :198-200)
  Mk5B frames   : 16-byte header (sync word 0xABADDEED little-endian, frame
                  number, BCD time) + 10000 payload bytes (FileSource.cpp:43-47)
"""
import numpy as np

LEVELS = np.array([1.0, -1.0, 3.3359, -3.3359], dtype=np.float32)


def quantize(x, sigma=1.0, thresh=0.98):
    """Real samples -> 2-bit fields (uint8 in 0..3)."""
    t = thresh * sigma
    f = np.where(x >= t, 2, np.where(x >= 0, 0, np.where(x >= -t, 1, 3)))
    return f.astype(np.uint8)


def pack_1ch(fields):
    f = np.asarray(fields, dtype=np.uint8).reshape(-1, 4)
    return (f[:, 0] << 6 | f[:, 1] << 4 | f[:, 2] << 2 | f[:, 3]).astype(np.uint8)


def pack_multi(fields_by_channel, increasing=True):
    """fields_by_channel: (nch, nsamples), nch multiple of 4 -> bytes."""
    f = np.asarray(fields_by_channel, dtype=np.uint8)
    nch, n = f.shape
    step = nch // 4
    out = np.zeros((n, step), dtype=np.uint8)
    for c in range(nch):
        if increasing:
            b, sh = c // 4, 2 * (c % 4)
        else:
            b, sh = (step - 1) - c // 4, 6 - 2 * (c % 4)
        out[:, b] |= (f[c] << sh).astype(np.uint8)
    return out.reshape(-1)


def mk5b_frames(payload, first_frame=0, seconds=0):
    """Wrap a payload (multiple of 10000 bytes) into Mk5B frames."""
    payload = np.asarray(payload, dtype=np.uint8)
    assert len(payload) % 10000 == 0
    nf = len(payload) // 10000
    out = np.zeros((nf, 10016), dtype=np.uint8)
    out[:, 16:] = payload.reshape(nf, 10000)
    hdr = np.zeros((nf, 4), dtype=np.uint32)
    hdr[:, 0] = 0xABADDEED
    hdr[:, 1] = (np.arange(nf) + first_frame) & 0x7FFF
    s = seconds
    bcd = 0
    for i in range(5):
        bcd |= (s % 10) << (4 * i)
        s //= 10
    hdr[:, 2] = bcd
    out[:, :16] = hdr.astype("<u4").view(np.uint8).reshape(nf, 16)
    return out.reshape(-1)


def gaussian_tones(rng, n, tones=(), sigma=1.0, start=0, fs=1.0):
    """sigma * N(0,1) + sum_k a_k cos(2 pi f_k (start+n)/fs + phi_k)."""
    x = rng.standard_normal(n) * sigma
    t = (np.arange(n, dtype=np.float64) + start) / fs
    for (a, f, ph) in tones:
        x += a * np.cos(2 * np.pi * f * t + ph)
    return x


def pcal_comb(fs, bw, offset=10e3, spacing=1e6, amp=0.141, rng=None):
    """Tone list of a phase-cal comb at offset + k*spacing below bw."""
    tones = []
    k = 0
    while offset + k * spacing < bw:
        ph = 0.0 if rng is None else float(rng.uniform(0, 2 * np.pi))
        tones.append((amp, offset + k * spacing, ph))
        k += 1
    return tones


def raw_1ch(seed, nsamples, tones=(), fs=32e6):
    rng = np.random.default_rng(seed)
    x = gaussian_tones(rng, nsamples, tones, fs=fs)
    return pack_1ch(quantize(x))


def raw_multi(seed, nch, nsamples, tones_per_ch=None, fs=32e6, increasing=True):
    rng = np.random.default_rng(seed)
    fields = []
    for c in range(nch):
        tn = tones_per_ch[c] if tones_per_ch else ()
        fields.append(quantize(gaussian_tones(rng, nsamples, tn, fs=fs)))
    return pack_multi(np.stack(fields), increasing=increasing)
