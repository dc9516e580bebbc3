"""Host-side mirror of the reference's caller of a compute core.

`TaskDispatcher` restates TaskDispatcher (src/TaskDispatcher.cpp:36-247): it
pre-fills one raw Buffer per core and source, then loops
    run every core on its current half -> read the other half ->
    join each core -> combineResults (spectrum spans several buffers) or write
until a source reports end of file.  Buffers are laid out as main() allocates
them (src/swspectrometer.cpp:365-413).  `FileSource` restates the raw-format
read of src/FileSource.cpp:166-250 (Mk5B frames: headers stripped on the
host, :184-215).  Output goes through `sinks.FileSink`, whose files are byte
for byte those of the reference's FileSink.
"""
from . import sinks as _sinks
from .core import Buffer, TaskCoreHIP

MK5B_FRAME, MK5B_HEADER, MK5B_PAYLOAD = 10016, 16, 10000


class FileSource:
    """src/FileSource.cpp:166-250.  read(buf) fills the whole allocation if it
    can and sets the Buffer length to the bytes read; eof() turns true once a
    read came up short (ifstream semantics, :235-240).  framed=True takes
    Mk5B frames and strips their 16-byte headers on the host; as in the
    reference (:184-215) a framed read always reports getAllocated() bytes."""

    def __init__(self, path, framed=False):
        self._f = open(path, "rb")
        self.framed = framed
        self._eof = False

    def read(self, buf):
        n = buf.getAllocated()
        dst = buf.getData()
        if not self.framed:
            got = self._f.readinto(memoryview(dst)[:n])
            if got < n:
                self._eof = True
            buf.setLength(got)
            return got
        # framed: payload bytes of successive frames, header skipped at each
        # frame start (sync word not checked here)
        off = 0
        while off < n and not self._eof:
            pos = self._f.tell()
            if pos % MK5B_FRAME == 0:
                if len(self._f.read(MK5B_HEADER)) < MK5B_HEADER:
                    self._eof = True
                continue
            avail = MK5B_FRAME - pos % MK5B_FRAME
            want = min(avail, n - off)
            got = self._f.readinto(memoryview(dst)[off: off + want])
            if got < avail:
                self._eof = True
            off += got
        buf.setLength(n)
        return n

    def eof(self):
        return self._eof

    def close(self):
        self._f.close()


class TaskDispatcher:
    """src/TaskDispatcher.cpp:36-247 with TaskCoreHIP cores.

    sources: one FileSource per input; sinks: num_sources auto sinks then
    num_xpols cross sinks; pcalsinks: one per source (phase-cal on).  cores
    are created and prepared here, `core_kwargs` going to prepare()."""

    def __init__(self, settings, sources, sinks, pcalsinks=(), num_cores=1, pinned=True, **core_kwargs):
        s = settings
        self.set = s
        self.sources, self.sinks, self.pcalsinks = list(sources), list(sinks), list(pcalsinks)
        self.ncores = num_cores
        nsrc, nx = s.num_sources, s.num_xpols
        # main(): raw Buffer[cores][2][sources], out[cores][sources], xpol[cores][xpols], pcal[cores][sources]
        self.raw = [[[Buffer(s.rawbuf_size, pinned=pinned) for _ in range(nsrc)] for _ in range(2)]
                    for _ in range(num_cores)]
        nsp = max(s.max_spectra_per_buffer, 1)
        self.out = [[Buffer(nsp * s.fft_bytes_ssb) for _ in range(nsrc)] for _ in range(num_cores)]
        self.xpol = [[Buffer(nsp * s.fft_bytes_xpol) for _ in range(nx)] for _ in range(num_cores)]
        self.pcal = [[Buffer(max(nsp * s.pcal_result_bytes, 1)) for _ in range(nsrc)] for _ in range(num_cores)]
        # prefill (:48-54)
        for c in range(num_cores):
            for i in range(nsrc):
                self.sources[i].read(self.raw[c][0][i])
        self.cores = []
        for c in range(num_cores):
            core = TaskCoreHIP(c)
            core.prepare(s, **core_kwargs)
            self.cores.append(core)
        # combined spectra and phase-cal (:68-82)
        peak = max(s.fft_bytes_xpol, s.fft_bytes_ssb)
        self.combined = [Buffer(peak) for _ in range(len(self.sinks))]
        for b in self.combined:
            b.setLength(peak)
            self.cores[0].resetBuffer(b)
        self.combined_pcal = [Buffer(max(s.pcal_result_bytes, 1)) for _ in range(nsrc)]
        for b in self.combined_pcal:
            b.setLength(s.pcal_result_bytes)
            self.cores[0].resetBuffer(b)
        self.total_spectra = 0
        self.total_corecompleted = 0

    def run(self):
        """Process all input (:92-247); returns the number of spectra written."""
        s = self.set
        nsrc, nx = s.num_sources, s.num_xpols
        dbl = 0
        num_combined = 0
        while True:
            for c, core in enumerate(self.cores):
                core.run(self.raw[c][dbl], self.out[c], self.xpol[c], self.pcal[c])
            for c in range(self.ncores):
                for i in range(nsrc):
                    self.sources[i].read(self.raw[c][(dbl + 1) % 2][i])
            for c, core in enumerate(self.cores):
                n = core.join()
                if n < 0:
                    raise RuntimeError("TaskCoreHIP %d join failed" % c)
                self.total_corecompleted += n
                if s.max_buffers_per_spectrum > 1:
                    core.combineResults(self.combined, self.combined_pcal)
                    num_combined += 1
                    if num_combined == s.max_buffers_per_spectrum:
                        for k, sk in enumerate(self.sinks):
                            sk.write(self.combined[k])
                            self.cores[0].resetBuffer(self.combined[k])
                        if s.extract_pcal:
                            for i in range(nsrc):
                                self.pcalsinks[i].write(self.combined_pcal[i])
                                self.cores[0].resetBuffer(self.combined_pcal[i])
                        self.total_spectra += 1
                        num_combined = 0
                else:
                    for k in range(min(len(self.sinks), nsrc)):
                        self.sinks[k].write(self.out[c][k])
                    for x in range(nx):
                        self.sinks[nsrc + x].write(self.xpol[c][x])
                    if s.extract_pcal:
                        for i in range(nsrc):
                            self.pcalsinks[i].write(self.pcal[c][i])
                            self.cores[0].resetBuffer(self.pcal[c][i])
                    self.total_spectra += s.max_spectra_per_buffer
            dbl = (dbl + 1) % 2
            # combined EOF check (:217-226): stop once a source is at EOF,
            # checked up to the first source whose next buffer holds data
            got_eof = False
            for i in range(nsrc):
                if self.sources[i].eof():
                    got_eof = True
                if self.raw[0][dbl][i].getLength() > 0:
                    break
            if got_eof:
                break
        for core in self.cores:
            core.finalize()
        for sk in self.sinks:
            sk.close()
        if s.extract_pcal:
            for sk in self.pcalsinks:
                sk.close()
        return self.total_spectra


def run_files(settings, inputs, out_prefix, fmt=_sinks.BINARY, framed=False, **core_kwargs):
    """Dispatch input files through TaskCoreHIP into files named
    <out_prefix>_s<i>.bin (auto), _x0.bin (cross) and _p<i>.bin (phase-cal);
    returns (spectra written, list of output paths)."""
    s = settings
    srcs = [FileSource(p, framed=framed) for p in inputs]
    paths = ["%s_s%d.bin" % (out_prefix, i) for i in range(s.num_sources)]
    paths += ["%s_x%d.bin" % (out_prefix, x) for x in range(s.num_xpols)]
    ppaths = ["%s_p%d.bin" % (out_prefix, i) for i in range(s.num_sources)] if s.extract_pcal else []
    sk = []
    for p in paths + ppaths:
        f = _sinks.FileSink(fmt)
        f.open(p)
        sk.append(f)
    d = TaskDispatcher(s, srcs, sk[: len(paths)], sk[len(paths):], **core_kwargs)
    n = d.run()
    for src in srcs:
        src.close()
    return n, paths + ppaths
