"""ctypes view of the CPU oracle (oracle/swspec_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (swspec_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswspec_oracle.so")

WINDOWS = {"None": 0, "Cosine": 1, "Cosine2": 2, "Hamming": 3, "Hann": 4, "Blackman": 5}
FORMATS = {"RawSigned": 0, "RawUnsigned": 1, "Mk5B": 2, "iBOB": 3, "VDIF": 4, "VLBA": 5,
           "MKIV": 6, "Mark5B": 7, "Maxim": 8}


class Ini(C.Structure):
    _fields_ = [
        ("num_cores", C.c_int), ("max_source_buffer_mb", C.c_double), ("platform_max_raw_mb", C.c_int),
        ("fft_points", C.c_long), ("fft_integ_seconds", C.c_float), ("fft_overlap_factor", C.c_int),
        ("window", C.c_int), ("bandwidth_hz", C.c_float), ("pcal_offset_hz", C.c_float),
        ("pcal_harmonics_hz", C.c_float), ("source_format", C.c_int), ("bits_per_sample", C.c_int),
        ("channelorder_increasing", C.c_int), ("source_channels", C.c_int),
        ("use_channel_file1", C.c_int), ("use_channel_file2", C.c_int), ("extract_pcal", C.c_int),
        ("num_sources", C.c_int), ("do_xpol", C.c_int), ("vlba_fanout", C.c_int),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("num_cores", C.c_int), ("fft_points", C.c_long), ("fft_integ_seconds", C.c_float),
        ("fft_overlap_factor", C.c_int), ("window", C.c_int), ("bits_per_sample", C.c_int),
        ("channelorder_increasing", C.c_int), ("source_channels", C.c_int),
        ("use_channel_file1", C.c_int), ("use_channel_file2", C.c_int),
        ("samplingfreq", C.c_float), ("pcaloffsethz", C.c_float), ("pcalharmonicshz", C.c_float),
        ("source_format", C.c_int), ("extract_pcal", C.c_int), ("num_sources", C.c_int),
        ("num_xpols", C.c_int), ("max_rawbuf_size", C.c_size_t), ("dt", C.c_float), ("df", C.c_float),
        ("averaged_ffts", C.c_int), ("fft_overlap_points", C.c_int), ("fft_ssb_points", C.c_int),
        ("rawbytes_per_channelsample", C.c_double), ("raw_fullfft_bytes", C.c_size_t),
        ("raw_overlap_bytes", C.c_size_t), ("fft_bytes", C.c_size_t), ("fft_bytes_ssb", C.c_size_t),
        ("fft_bytes_xpol", C.c_size_t), ("rawbuf_size", C.c_size_t), ("max_spectra_per_buffer", C.c_int),
        ("max_specffts_per_buffer", C.c_int), ("max_buffers_per_spectrum", C.c_int),
        ("core_overlapped_ffts", C.c_int), ("core_averaged_ffts", C.c_int), ("pcal_tonebins", C.c_int),
        ("pcal_rotatorlen", C.c_int), ("pcal_pulses", C.c_int), ("pcal_pulses_per_fft", C.c_int),
        ("pcal_result_bytes", C.c_size_t), ("vlba_fanout", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    """Compile the oracle with gcc (a checker, not product code)."""
    subprocess.check_call(["make", "-s", "-C", HERE])


REF_SRC = "/root/reference/src"
REF_FILESINK = os.path.join(HERE, "_ref", "filesink_ref")
REF_ADAPTER_DRV = os.path.join(HERE, "_ref", "taskcorehip_drv")


def build_ref():
    """Compile oracle/_ref/ from the reference's own sources when they are
    present: the reference FileSink and Buffer with our sink driver, and the
    C++ TaskCoreHIP adapter driver (needs swspec_amd/libswspec_hip.so built
    first).  Returns whether the reference FileSink binary is available."""
    if os.path.isdir(REF_SRC):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref", "REFSRC=" + REF_SRC])
        if os.path.exists(os.path.join(HERE, "..", "swspec_amd", "libswspec_hip.so")):
            subprocess.check_call(["make", "-s", "-C", HERE, "adapter", "REFSRC=" + REF_SRC])
    return os.path.exists(REF_FILESINK)


def ref_filesink(fmt, arrays, out_path, lengths=None):
    """Write float32 arrays through the reference's FileSink::write (one
    Buffer per array, length = its bytes unless `lengths` gives them);
    returns the per-write return values."""
    import tempfile
    bufs = [np.ascontiguousarray(a, dtype=np.float32).view(np.uint8).ravel() for a in arrays]
    lens = lengths if lengths is not None else [b.size for b in bufs]
    # the driver takes each write's bytes in turn from one input file
    data = np.concatenate([b[:int(n)] for b, n in zip(bufs, lens)]) if bufs else np.zeros(0, np.uint8)
    with tempfile.NamedTemporaryFile(suffix=".f32", delete=False) as t:
        data.tofile(t.name)
    try:
        out = subprocess.run([REF_FILESINK, fmt, t.name, str(out_path)] + [str(int(x)) for x in lens],
                             capture_output=True, text=True, check=True).stdout
    finally:
        os.unlink(t.name)
    return [int(v) for v in out.split()]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.so_derive.argtypes = [C.POINTER(Ini), C.POINTER(Settings)]
        L.so_gcd.restype = C.c_longlong
        L.so_gcd.argtypes = [C.c_longlong, C.c_longlong]
        L.so_pick_unpacker.argtypes = [C.POINTER(Settings)]
        L.so_unpack.restype = C.c_size_t
        L.so_unpack.argtypes = [C.POINTER(Settings), P, P, C.c_size_t, C.c_int]
        L.so_window_table.argtypes = [C.c_int, C.c_long, P]
        L.so_window_table64.argtypes = [C.c_int, C.c_long, P]
        L.so_rfft64.argtypes = [P, C.c_long, P]
        L.so_rfft32.argtypes = [P, C.c_long, P]
        L.so_cfft64.argtypes = [P, C.c_long]
        L.so_count_spectra.argtypes = [C.POINTER(Settings), C.c_size_t, C.POINTER(C.c_int)]
        L.so_domaths.argtypes = [C.POINTER(Settings), C.c_int, C.POINTER(P), C.POINTER(C.c_size_t),
                                 C.POINTER(P), C.POINTER(P), C.POINTER(P)]
        L.so_extract_pcal32.argtypes = [C.POINTER(Settings), P, C.c_long, P]
        L.so_extract_pcal64.argtypes = [C.POINTER(Settings), P, C.c_long, P]
        L.so_pcal_fold64.argtypes = [P, C.c_long, C.c_double, C.c_double, C.c_int, C.c_long,
                                     C.c_uint64, C.c_int, P]
        L.so_pcal_tones64.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_double, P]
        L.so_pcal_ntones.argtypes = [C.c_double, C.c_double, C.c_double]
        L.so_pcal_reference64.argtypes = [P, C.c_long, C.c_double, C.c_double, C.c_double, C.c_uint64, P]
        L.so_pcalx_init.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_int, C.c_size_t]
        L.so_pcalx_free.argtypes = [P]
        L.so_pcalx_clear.argtypes = [P]
        L.so_pcalx_adjust.argtypes = [P, C.c_size_t]
        L.so_pcalx_extract.argtypes = [P, P, C.c_size_t]
        L.so_pcalx_final.restype = C.c_uint64
        L.so_pcalx_final.argtypes = [P, P]
        L.so_pcal_delay64.restype = C.c_double
        L.so_pcal_delay64.argtypes = [P, C.c_int, C.c_double, C.c_int]
        L.so_mk5b_strip.restype = C.c_long
        L.so_mk5b_strip.argtypes = [P, C.c_long, P]
        L.so_vlba_modulate.argtypes = [P]
        L.so_vlba_modulate.restype = None
        L.so_vlba_blank.argtypes = [P, C.c_long, C.c_long, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.so_vlba_blank.restype = None
        L.so_vlba_unpack.argtypes = [P, C.c_long, C.c_size_t, C.c_int, P]
        L.so_vlba_unpack.restype = C.c_size_t
        L.so_vlba_unpack_mode.argtypes = [P, C.c_long, C.c_size_t, C.c_int, C.c_int, C.c_int, P]
        L.so_vlba_unpack_mode.restype = C.c_size_t
        L.so_vlba_blank_zone.argtypes = [P, C.c_long, C.c_long, C.c_int, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
        L.so_vlba_blank_zone.restype = None
        L.so_mark5b_unpack.argtypes = [P, C.c_long, C.c_size_t, C.c_int, C.c_int, C.c_int, P]
        L.so_mark5b_unpack.restype = C.c_size_t
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def make_ini(**kw):
    """Defaults of src/swspectrometer.cpp:67-91 (platform INTEL_IPP: 32 MB)."""
    d = dict(num_cores=1, max_source_buffer_mb=-1.0, platform_max_raw_mb=32, fft_points=320000,
             fft_integ_seconds=20.0, fft_overlap_factor=2, window=WINDOWS["Cosine2"], bandwidth_hz=8e6,
             pcal_offset_hz=10e3, pcal_harmonics_hz=1e6, source_format=FORMATS["RawSigned"],
             bits_per_sample=8, channelorder_increasing=1, source_channels=1, use_channel_file1=1,
             use_channel_file2=1, extract_pcal=0, num_sources=1, do_xpol=0, vlba_fanout=1)
    for k, v in kw.items():
        if k == "window" and isinstance(v, str):
            v = WINDOWS[v]
        if k == "source_format" and isinstance(v, str):
            # FileSource::open (src/FileSource.cpp:41-106): mark5access names by prefix;
            # "VLBA1_<fanout>-<Mbps>-<nchan>-<nbit>" carries the fanout
            if v.upper().startswith("VLBA1_"):
                d["vlba_fanout"] = int(v[6:].split("-")[0])
            if v not in FORMATS:
                v = next((FORMATS[p] for p in ("MKIV", "VLBA", "Mark5B") if v.lower().startswith(p.lower())), -1)
            else:
                v = FORMATS[v]
        d[k] = v
    ini = Ini()
    for k, v in d.items():
        setattr(ini, k, v)
    return ini


def derive(**kw):
    ini = make_ini(**kw)
    s = Settings()
    rc = lib().so_derive(C.byref(ini), C.byref(s))
    if rc != 0:
        raise ValueError("so_derive failed rc=%d" % rc)
    return s


def gcd(a, b):
    return lib().so_gcd(a, b)


def unpack(s, raw, count, channel):
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    out = np.zeros(count + 32, dtype=np.float32)
    n = lib().so_unpack(C.byref(s), _p(raw), _p(out), count, channel)
    return out[:n]


def window(kind, n, fp64=False):
    if isinstance(kind, str):
        kind = WINDOWS[kind]
    if fp64:
        out = np.empty(n, dtype=np.float64)
        lib().so_window_table64(kind, n, _p(out))
    else:
        out = np.empty(n, dtype=np.float32)
        lib().so_window_table(kind, n, _p(out))
    return out


def rfft64(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    X = np.empty(2 * (len(x) // 2 + 1), dtype=np.float64)
    if lib().so_rfft64(_p(x), len(x), _p(X)) != 0:
        raise ValueError("unsupported length")
    return X[0::2] + 1j * X[1::2]


def rfft32(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    X = np.empty(2 * (len(x) // 2 + 1), dtype=np.float32)
    if lib().so_rfft32(_p(x), len(x), _p(X)) != 0:
        raise ValueError("unsupported length")
    return (X[0::2] + 1j * X[1::2]).astype(np.complex64)


def cfft64(z):
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    if lib().so_cfft64(_p(z), len(z)) != 0:
        raise ValueError("unsupported length")
    return z


def count_spectra(s, nbytes):
    n = C.c_int(0)
    sp = lib().so_count_spectra(C.byref(s), nbytes, C.byref(n))
    return sp, n.value


def domaths(s, raws, precision=64):
    """doMaths on one buffer per source; returns (nspectra, auto, xpol, pcal)."""
    nsrc = s.num_sources
    raws = [np.ascontiguousarray(r, dtype=np.uint8) for r in raws]
    nbytes = min(len(r) for r in raws)
    nsp, _ = count_spectra(s, nbytes)
    nsp_alloc = max(nsp, 1)
    ssb = s.fft_ssb_points
    ft = np.float32 if precision == 32 else np.float64
    auto = [np.zeros(nsp_alloc * ssb, dtype=ft) for _ in range(nsrc)]
    xpol = [np.zeros(2 * nsp_alloc * ssb, dtype=ft) for _ in range(s.num_xpols)]
    T = max(s.pcal_tonebins, 1)
    pcal = [np.zeros(2 * nsp_alloc * T, dtype=ft) for _ in range(nsrc)]
    P = C.c_void_p
    srcs = (P * 2)(*[_p(r) for r in raws] + [None] * (2 - nsrc))
    lens = (C.c_size_t * 2)(*[len(r) for r in raws] + [0] * (2 - nsrc))
    oa = (P * 2)(*[_p(a) for a in auto] + [None] * (2 - nsrc))
    ox = (P * 1)(*([_p(xpol[0])] if xpol else [None]))
    op = (P * 2)(*[_p(a) for a in pcal] + [None] * (2 - nsrc))
    n = lib().so_domaths(C.byref(s), precision, srcs, lens, oa, ox, op)
    auto = [a.reshape(nsp_alloc, ssb)[:n] for a in auto]
    xpol = [(x[0::2] + 1j * x[1::2]).reshape(nsp_alloc, ssb)[:n] for x in xpol]
    pcal = [(p[0::2] + 1j * p[1::2]).reshape(nsp_alloc, T)[:n] for p in pcal]
    return n, auto, xpol, pcal


def pcal_fold64(x, fs_hz, offset_hz, nbins, rotatorlen, sampleoffset=0, rot32=False, acc=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    if acc is None:
        acc = np.zeros(2 * nbins, dtype=np.float64)
    lib().so_pcal_fold64(_p(x), len(x), fs_hz, offset_hz, nbins, rotatorlen, sampleoffset,
                         1 if rot32 else 0, _p(acc))
    return acc


def pcal_tones64(folded, ntones, spacing_hz, fs_hz):
    f = np.ascontiguousarray(folded)
    if np.iscomplexobj(f):
        f = np.ascontiguousarray(np.stack([f.real, f.imag], axis=-1).reshape(-1), dtype=np.float64)
    nbins = len(f) // 2
    out = np.zeros(2 * ntones, dtype=np.float64)
    lib().so_pcal_tones64(_p(f), nbins, ntones, spacing_hz, fs_hz, _p(out))
    return out[0::2] + 1j * out[1::2]


def pcal_ntones(bw, spacing, offset):
    return lib().so_pcal_ntones(bw, spacing, offset)


def pcal_reference64(x, bw, spacing, offset, sampleoffset=0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    nt = pcal_ntones(bw, spacing, offset)
    out = np.zeros(2 * nt, dtype=np.float64)
    lib().so_pcal_reference64(_p(x), len(x), bw, spacing, offset, sampleoffset, _p(out))
    return out[0::2] + 1j * out[1::2]


class PCalX(C.Structure):
    """so_pcalx: the PCal extractor classes of PCal.cpp (FP64 accumulation)."""
    _fields_ = [("kind", C.c_int), ("fs_hz", C.c_double), ("spacing_hz", C.c_double), ("offset_hz", C.c_int),
                ("nbins", C.c_int), ("ntones", C.c_int), ("rotatorlen", C.c_long), ("pcal_index", C.c_size_t),
                ("rotator_index", C.c_size_t), ("samplecount", C.c_uint64), ("finalized", C.c_int),
                ("buf", C.c_void_p), ("rot", C.c_void_p), ("dft", C.c_void_p)]


PCAL_KINDS = {"Trivial": 0, "Shifting": 1, "ImplicitShift": 2}


class PCalExtractor:
    """PCal::getNew (kind=None) or one class by name; methods as PCal.h."""

    def __init__(self, bandwidth_hz, spacing_hz, offset_hz, sampleoffset=0, kind=None):
        self.s = PCalX()
        k = -1 if kind is None else PCAL_KINDS[kind]
        if lib().so_pcalx_init(C.byref(self.s), k, bandwidth_hz, spacing_hz, int(offset_hz), sampleoffset) != 0:
            raise ValueError("bad PCal parameters")
        self.kind = {v: k for k, v in PCAL_KINDS.items()}[self.s.kind]

    def __del__(self):
        if getattr(self, "s", None) is not None and _lib is not None:
            _lib.so_pcalx_free(C.byref(self.s))
            self.s = None

    def clear(self):
        lib().so_pcalx_clear(C.byref(self.s))

    def adjustSampleOffset(self, so):
        lib().so_pcalx_adjust(C.byref(self.s), so)

    def extractAndIntegrate(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        return bool(lib().so_pcalx_extract(C.byref(self.s), _p(x), len(x)))

    def getLength(self):
        return self.s.ntones

    def getNBins(self):
        return self.s.nbins

    def getFinalPCal(self, out=None):
        """Returns (samplecount, tones complex128); `out` seeds the entries
        past the selection's break (left as they were)."""
        o = np.zeros(2 * self.s.ntones) if out is None else np.stack([out.real, out.imag], -1).reshape(-1).copy()
        n = lib().so_pcalx_final(C.byref(self.s), _p(o))
        return int(n), o[0::2] + 1j * o[1::2]


def pcal_delay64(tones, spacing_hz, upsample=64):
    t = np.ascontiguousarray(np.stack([np.real(tones), np.imag(tones)], -1).reshape(-1), dtype=np.float64)
    return lib().so_pcal_delay64(_p(t), len(tones), spacing_hz, upsample)


def mk5b_strip(frames):
    frames = np.ascontiguousarray(frames, dtype=np.uint8)
    out = np.zeros(len(frames), dtype=np.uint8)
    n = lib().so_mk5b_strip(_p(frames), len(frames), _p(out))
    if n < 0:
        raise ValueError("Mk5B sync word missing")
    return out[:n]


VLBA_PAYLOAD = 20000
VLBA_CHUNK = 40128


def vlba_modulate():
    out = np.zeros(VLBA_PAYLOAD, dtype=np.uint8)
    lib().so_vlba_modulate(_p(out))
    return out


def vlba_blank(buf, seg, avail=None):
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    vs, ve = C.c_int(), C.c_int()
    lib().so_vlba_blank(_p(buf), len(buf) if avail is None else avail, seg, C.byref(vs), C.byref(ve))
    return vs.value, ve.value


def vlba_unpack(raw, count, channel, avail=None, fanout=1, nbit=2):
    """VLBAUnpacker::extract_samples for the one-byte-per-time-sample modes
    VLBA1_<fanout>-*-4-2 (nbit 2) and VLBA1_<fanout>-*-8-1 (nbit 1)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    out = np.zeros(count, dtype=np.float32)
    n = lib().so_vlba_unpack_mode(_p(raw), len(raw) if avail is None else avail, count, channel, fanout, nbit,
                                  _p(out))
    if n != count:
        raise ValueError("VLBA mode fanout %d nbit %d not restated" % (fanout, nbit))
    return out


def mark5b_unpack(raw, count, channel, nchan, nbit, avail=None):
    """Mark5BUnpacker::extract_samples over mark5access's Mark5B decoders
    (samples 0..15: the fixed +-3.3359 pattern that stands in for std::rand())."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    out = np.zeros(count, dtype=np.float32)
    n = lib().so_mark5b_unpack(_p(raw), len(raw) if avail is None else avail, count, channel, nchan, nbit, _p(out))
    if n != count:
        raise ValueError("Mark5B configuration %d channels x %d bits not restated" % (nchan, nbit))
    return out


def vlba_blank_zone(buf, seg, payload, zone, avail=None):
    """Blank zone `zone` of the `payload`-byte payload at byte `seg` (payload-relative bounds)."""
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    vs, ve = C.c_int(), C.c_int()
    lib().so_vlba_blank_zone(_p(buf), len(buf) if avail is None else avail, seg, payload, zone, C.byref(vs),
                             C.byref(ve))
    return vs.value, ve.value
