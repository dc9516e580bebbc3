"""Host-side mirror of the reference's compute-core plugin interface.

`TaskCoreHIP` has the methods of class TaskCore (src/TaskCore.h:36-89) with
the same argument meaning and return values, backed by the C ABI of
libswspec_hip.so.  `Buffer` mirrors src/Buffer.h:27-70 and `derive_settings`
restates the INI -> settings derivation of src/swspectrometer.cpp:67-231
(implemented in the library, swspec_hip_derive).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FORMATS, WINDOWS, Cfg, Derived, Ini, check, load

# Reference defaults (src/swspectrometer.cpp:67-91); PLATFORM_MAX_RAW_BUF_SIZE_MB
# of the IPP build (src/Settings.h:30-31).
INI_DEFAULTS = dict(
    NumCPUCores=1, MaxSourceBufferMB=None, FFTpoints=320000, FFTIntegrationTimeSec=20.0,
    FFToverlapFactor=2, WindowType="Cosine2", BandwidthHz=8e6, PCalOffsetHz=10e3, SourceFormat="RawSigned",
    BitsPerSample=8, ChannelOrderIncreasing=True, SourceChannels=1, UseFile1Channel=1, UseFile2Channel=1,
    ExtractPCal=False, DoCrossPolarization=False,
)
PLATFORM_MAX_RAW_BUF_SIZE_MB = 32
PCAL_HARMONICS_HZ = 1e6


class Settings:
    """Derived settings (the fields of swspect_settings_t a core reads)."""

    def __init__(self, keys, ini, derived, num_sources):
        self.keys = keys
        self.ini = ini
        self.derived = derived
        self.num_sources = num_sources
        for k, v in derived.as_dict().items():
            setattr(self, k, v)
        self.fft_points = ini.fft_points
        self.fft_overlap_factor = ini.fft_overlap_factor
        self.num_cores = ini.num_cores
        self.fft_bytes_ssb = 4 * self.fft_ssb_points
        self.fft_bytes_xpol = 2 * self.fft_bytes_ssb
        self.pcal_result_bytes = self.pcal_tonebins * 8 if self.extract_pcal else 0

    def cfg(self, device=0, flags=0, ffts_per_chunk=0):
        c = Cfg()
        check(load().swspec_hip_cfg_from_ini(C.byref(self.ini), C.byref(self.derived), C.byref(c)),
              what="swspec_hip_cfg_from_ini")
        c.device = device
        c.flags = flags
        c.ffts_per_chunk = ffts_per_chunk
        f = self.keys["SourceFormat"]
        if isinstance(f, str):
            c.mark5_format = f.encode()[:31]
        return c


def source_format_enum(name):
    """FileSource::open's format selection (src/FileSource.cpp:41-106): exact
    names, or the mark5access prefixes MKIV / VLBA / Mark5B (case-insensitive)."""
    if not isinstance(name, str):
        return int(name)
    low = name.lower()
    for k, v in FORMATS.items():
        if k.lower() == low:
            return v
    for prefix in ("MKIV", "VLBA", "Mark5B"):
        if low.startswith(prefix.lower()):
            return FORMATS[prefix]
    return -1


def derive_settings(num_sources=1, **keys):
    """INI [Spectrometer] keys -> Settings (src/swspectrometer.cpp:101-231)."""
    k = dict(INI_DEFAULTS)
    k.update(keys)
    ini = Ini()
    ini.num_cores = int(k["NumCPUCores"])
    ini.max_source_buffer_mb = -1.0 if k["MaxSourceBufferMB"] is None else float(k["MaxSourceBufferMB"])
    ini.platform_max_raw_mb = PLATFORM_MAX_RAW_BUF_SIZE_MB
    ini.fft_points = int(k["FFTpoints"])
    ini.fft_integ_seconds = float(k["FFTIntegrationTimeSec"])
    ini.fft_overlap_factor = int(k["FFToverlapFactor"])
    w = k["WindowType"]
    ini.window = WINDOWS.get(w, WINDOWS["Cosine2"]) if isinstance(w, str) else int(w)
    ini.bandwidth_hz = float(k["BandwidthHz"])
    ini.pcal_offset_hz = float(k["PCalOffsetHz"])
    ini.pcal_harmonics_hz = PCAL_HARMONICS_HZ
    ini.source_format = source_format_enum(k["SourceFormat"])
    ini.bits_per_sample = int(k["BitsPerSample"])
    ini.channelorder_increasing = 1 if k["ChannelOrderIncreasing"] else 0
    ini.source_channels = int(k["SourceChannels"])
    ini.use_channel_file1 = int(k["UseFile1Channel"])
    ini.use_channel_file2 = int(k["UseFile2Channel"])
    ini.extract_pcal = 1 if k["ExtractPCal"] else 0
    ini.num_sources = num_sources
    ini.do_xpol = 1 if (k["DoCrossPolarization"] and num_sources == 2) else 0
    d = Derived()
    check(load().swspec_hip_derive(C.byref(ini), C.byref(d)), what="swspec_hip_derive")
    return Settings(k, ini, d, num_sources)


class Buffer:
    """src/Buffer.h:27-70: an aligned byte buffer with a length field.

    With pinned=True the bytes come from hipHostMalloc so that the core's
    host<->device copies run asynchronously."""

    def __init__(self, nbytes, pinned=False):
        self._pinned = None
        if pinned:
            p = load().swspec_hip_host_alloc(nbytes)
            if not p:
                raise MemoryError("hipHostMalloc(%d) failed" % nbytes)
            self._pinned = p
            self.data = np.ctypeslib.as_array((C.c_uint8 * nbytes).from_address(p))
        else:
            self.data = np.zeros(nbytes, dtype=np.uint8)
        self.length = 0

    def __del__(self):
        if self._pinned:
            try:
                load().swspec_hip_host_free(self._pinned)
            except Exception:
                pass
            self._pinned = None

    def getData(self):
        return self.data

    def getLength(self):
        return self.length

    def getAllocated(self):
        return self.data.nbytes

    def setLength(self, n):
        # src/Buffer.cpp:91-99: crop to the allocation
        self.length = min(int(n), self.getAllocated())

    def ptr(self):
        return self.data.ctypes.data


class TaskCoreHIP:
    """TaskCore (src/TaskCore.h:36-89) on the MI355X core."""

    def __init__(self, rank=0):
        self.rank = rank
        self._core = None
        self._running = False

    # TaskCore::prepare
    def prepare(self, settings, device=0, flags=0, ffts_per_chunk=0, channels=None):
        self.cfg_settings = settings
        cfg = settings.cfg(device=device, flags=flags, ffts_per_chunk=ffts_per_chunk)
        core = C.c_void_p()
        if channels is None:
            rc = load().swspec_hip_create(C.byref(cfg), self.rank, C.byref(core))
            self.njobs = settings.num_sources
        else:
            arr = (C.c_int * len(channels))(*channels)
            rc = load().swspec_hip_create_multichannel(C.byref(cfg), len(channels), arr, self.rank,
                                                       C.byref(core))
            self.njobs = len(channels)
        check(rc, None, "swspec_hip_create")
        self._core = core
        self.cfg = cfg

    def _pp(self, ptrs):
        return (C.c_void_p * max(len(ptrs), 1))(*ptrs)

    # TaskCore::run (host Buffers)
    def run(self, inbuf, outbuf, xpolbuf, pcalbuf):
        s = self.cfg_settings
        ins = self._pp([b.ptr() for b in inbuf])
        lens = (C.c_size_t * len(inbuf))(*[b.getLength() for b in inbuf])
        outs = self._pp([b.ptr() for b in outbuf[: self.njobs]])
        xp = self._pp([xpolbuf[0].ptr()] if (xpolbuf and s.num_xpols) else [None])
        pc = self._pp([b.ptr() for b in pcalbuf[: self.njobs]] if (pcalbuf and s.extract_pcal) else [None] * self.njobs)
        self._bufs = (outbuf, xpolbuf, pcalbuf)
        check(load().swspec_hip_run(self._core, ins, lens, outs, xp, pc), self._core, "swspec_hip_run")
        self._running = True
        return 0

    # device-resident variant (torch tensors or raw device pointers)
    def run_device(self, d_in, in_len, d_auto, d_xpol=None, d_pcal=None):
        def ptr(x):
            return x if isinstance(x, int) else x.data_ptr()
        ins = self._pp([ptr(x) for x in d_in])
        lens = (C.c_size_t * len(d_in))(*in_len)
        outs = self._pp([ptr(x) for x in d_auto])
        xp = self._pp([ptr(d_xpol[0])] if d_xpol else [None])
        pc = self._pp([ptr(x) for x in d_pcal] if d_pcal else [None] * self.njobs)
        self._bufs = None
        check(load().swspec_hip_run_device(self._core, ins, lens, outs, xp, pc), self._core,
              "swspec_hip_run_device")
        self._running = True
        return 0

    # TaskCore::join
    def join(self):
        n = C.c_int(0)
        la, lx, lp = C.c_size_t(0), C.c_size_t(0), C.c_size_t(0)
        rc = load().swspec_hip_join(self._core, C.byref(n), C.byref(la), C.byref(lx), C.byref(lp))
        self._running = False
        if rc != 0:
            return -1
        if self._bufs is not None:
            outbuf, xpolbuf, pcalbuf = self._bufs
            for b in outbuf[: self.njobs]:
                b.setLength(la.value)
            if xpolbuf and self.cfg_settings.num_xpols:
                xpolbuf[0].setLength(lx.value)
            if pcalbuf and self.cfg_settings.extract_pcal:
                for b in pcalbuf[: self.njobs]:
                    b.setLength(lp.value)
        return n.value

    # TaskCore::combineResults
    def combineResults(self, outbuf, pcalbuf):
        outs = self._pp([b.ptr() for b in outbuf])
        pc = self._pp([b.ptr() for b in pcalbuf] if pcalbuf else [None] * self.njobs)
        ol = (C.c_size_t * len(outbuf))()
        pl = (C.c_size_t * max(len(pcalbuf) if pcalbuf else 1, 1))()
        check(load().swspec_hip_combine(self._core, outs, pc, ol, pl), self._core, "swspec_hip_combine")
        for i, b in enumerate(outbuf):
            if ol[i]:
                b.setLength(ol[i])
        if pcalbuf:
            for i, b in enumerate(pcalbuf):
                if pl[i]:
                    b.setLength(pl[i])
        return 0

    # TaskCore::resetBuffer
    def resetBuffer(self, buf):
        load().swspec_hip_reset(buf.ptr(), buf.getAllocated())

    # combineResults for a run_device run: device combined arrays (tensors or
    # pointers), num_sources auto arrays then num_xpols cross arrays; returns
    # (out lengths, pcal lengths) in bytes as combineResults sets them
    def combine_device(self, d_out, d_pcal=None):
        def ptr(x):
            return x if isinstance(x, int) else x.data_ptr()
        outs = self._pp([ptr(x) for x in d_out])
        pc = self._pp([ptr(x) for x in d_pcal] if d_pcal else [None] * self.njobs)
        ol = (C.c_size_t * max(len(d_out), 1))()
        pl = (C.c_size_t * max(len(d_pcal) if d_pcal else 1, 1))()
        check(load().swspec_hip_combine_device(self._core, outs, pc, ol, pl), self._core,
              "swspec_hip_combine_device")
        return list(ol)[: len(d_out)], (list(pl)[: len(d_pcal)] if d_pcal else [])

    # resetBuffer for a device array (tensor or pointer) of `nbytes`
    @staticmethod
    def reset_device(d_buf, nbytes=None):
        p = d_buf if isinstance(d_buf, int) else d_buf.data_ptr()
        n = nbytes if nbytes is not None else d_buf.numel() * d_buf.element_size()
        check(load().swspec_hip_reset_device(p, n), None, "swspec_hip_reset_device")

    # TaskCore::finalize
    def finalize(self):
        if self._core is not None:
            check(load().swspec_hip_finalize(self._core), None, "swspec_hip_finalize")
            self._core = None
        return 0

    def pcal_tones(self, source, spectrum):
        nt = load().swspec_hip_pcal_ntones(self._core)
        out = np.zeros(2 * max(nt, 1), dtype=np.float64)
        check(load().swspec_hip_pcal_tones(self._core, source, spectrum, out.ctypes.data, nt), self._core,
              "swspec_hip_pcal_tones")
        return out[0: 2 * nt: 2] + 1j * out[1: 2 * nt: 2]

    def set_profiling(self, on=True, kernels=None):
        """HIP events around each launch; kernels: names of kernel_stats to
        time (default every kernel)."""
        v = 1 if on else 0
        if on and kernels is not None:
            names = ["pass1", "pass2", "pcal", "finalize"]
            v = 0x100
            for k in kernels:
                v |= 1 << names.index(k)
        check(load().swspec_hip_set_profiling(self._core, v), self._core, "swspec_hip_set_profiling")

    def kernel_stats(self):
        """{name: (total_ms, launches)} from HIP events around each launch."""
        out = {}
        for k, name in enumerate(["pass1", "pass2", "pcal", "finalize"]):
            ms, n = C.c_double(0), C.c_int64(0)
            check(load().swspec_hip_kernel_stats(self._core, k, C.byref(ms), C.byref(n)), self._core,
                  "swspec_hip_kernel_stats")
            out[name] = (ms.value, n.value)
        return out

    def last_run_seconds(self):
        return load().swspec_hip_last_run_seconds(self._core)

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass


def unpack_device(settings, d_raw, channel, count, d_out, device=0, flags=0):
    """DataUnpacker::extract_samples on the GPU (device pointers or tensors).
    flags=FLAG_MK5B_FRAMED: d_raw holds whole Mk5B frames (headers included)."""
    cfg = settings.cfg(device=device, flags=flags)
    raw = d_raw if isinstance(d_raw, int) else d_raw.data_ptr()
    out = d_out if isinstance(d_out, int) else d_out.data_ptr()
    check(load().swspec_hip_unpack_device(C.byref(cfg), raw, channel, count, out), None,
          "swspec_hip_unpack_device")


def plan_info(fft_points, flags=0):
    info = (C.c_int32 * 16)()
    check(load().swspec_hip_plan_info(fft_points, flags, info, 16), None, "swspec_hip_plan_info")
    keys = ["N1", "N2", "B", "p1_threads", "p2_threads", "npairs", "p1_maxpts", "p2_maxpts",
            "p1_cx", "p2_cx", "p1_stages", "p2_stages", "mfma64", "p1_radix0", "single", "p2_pad"]
    return dict(zip(keys, list(info)))


def autotune(fft_points, window="Cosine2", nfft=8, device=0):
    """Measure the planner's distinct plans for fft_points on the device and
    keep the fastest for every later plan of that length (cores, server jobs,
    plan_info).  Returns the chosen knobs and the times (ms per pass 1 +
    pass 2 over nfft FFTs) of the chosen and of the planner's own plan."""
    from ._lib import WINDOWS
    w = WINDOWS[window] if isinstance(window, str) else int(window)
    k = (C.c_int32 * 4)()
    best, dflt, n = C.c_double(0), C.c_double(0), C.c_int32(0)
    check(load().swspec_hip_autotune(device, int(fft_points), w, int(nfft), k, C.byref(best), C.byref(dflt),
                                     C.byref(n)), None, "swspec_hip_autotune")
    return {"p1_maxpts": k[0], "p1_bmax": k[1], "cx_mask": k[2], "p2_maxpts": k[3], "ms_best": best.value,
            "ms_default": dflt.value, "candidates": n.value}


def autotune_clear():
    """Forget every autotuned choice."""
    load().swspec_hip_autotune_clear()
