"""Host logic of the dispatcher mirror (no GPU): FileSource reads as the
reference's FileSource::read (src/FileSource.cpp:166-250) -- raw formats fill
the allocation and turn eof() on at the first short read; Mk5B frames have
their 16-byte headers stripped and the read always reports the allocation --
checked against the oracle's header strip."""
import numpy as np

from swspec_amd import synth
from swspec_amd.dispatcher import FileSource


class _Buf:
    def __init__(self, n):
        self.data = np.zeros(n, dtype=np.uint8)
        self.length = -1

    def getData(self):
        return self.data

    def getAllocated(self):
        return self.data.nbytes

    def setLength(self, n):
        self.length = int(n)


def test_raw_reads_and_eof(tmp_path):
    data = np.random.default_rng(1).integers(0, 256, 2500, dtype=np.uint8)
    p = tmp_path / "raw.bin"
    data.tofile(str(p))
    src = FileSource(str(p))
    got = []
    for _ in range(3):
        b = _Buf(1000)
        src.read(b)
        got.append((b.length, src.eof()))
        if b.length > 0:
            assert np.array_equal(b.data[: b.length], data[1000 * (len(got) - 1): 1000 * (len(got) - 1) + b.length])
    assert got == [(1000, False), (1000, False), (500, True)]
    # an exact multiple: eof only after a read that returns nothing
    p2 = tmp_path / "raw2.bin"
    data[:2000].tofile(str(p2))
    src = FileSource(str(p2))
    lens = []
    for _ in range(3):
        b = _Buf(1000)
        src.read(b)
        lens.append((b.length, src.eof()))
    assert lens == [(1000, False), (1000, False), (0, True)]


def test_framed_reads_strip_headers(tmp_path, oracle):
    payload = np.random.default_rng(2).integers(0, 256, 5 * 10000, dtype=np.uint8)
    frames = synth.mk5b_frames(payload)
    p = tmp_path / "frames.bin"
    frames.tofile(str(p))
    src = FileSource(str(p), framed=True)
    out = []
    for _ in range(2):
        b = _Buf(20000)
        n = src.read(b)
        assert n == 20000 and b.length == 20000
        out.append(b.data.copy())
    assert np.array_equal(np.concatenate(out), oracle.mk5b_strip(frames)[:40000])
    b = _Buf(20000)
    assert src.read(b) == 20000          # the reference reports the allocation even at the end
    assert np.array_equal(b.data[:10000], payload[40000:])
    assert src.eof()
