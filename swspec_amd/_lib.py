"""ctypes binding of libswspec_hip.so (the C ABI declared in include/swspec_hip.h).

The library is built in-tree by swspec_amd.build.build(); importing a product
path without it raises, there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswspec_hip.so")

SWSPEC_HIP_ABI_VERSION = 2
FLAG_NO_KAHAN = 0x1
FLAG_MK5B_FRAMED = 0x2   # Mk5B input as whole 10016-byte frames; headers skipped on the device
FLAG_MFMA64 = 0x4        # first column stage: radix 64 as a split-FP16 MFMA product

WINDOWS = {"None": 0, "Cosine": 1, "Cosine2": 2, "Hamming": 3, "Hann": 4, "Blackman": 5}
FORMATS = {"RawSigned": 0, "RawUnsigned": 1, "Mk5B": 2, "iBOB": 3, "VDIF": 4, "VLBA": 5,
           "MKIV": 6, "Mark5B": 7, "Maxim": 8}

# Every symbol include/swspec_hip.h declares (checked by the CPU tests).
EXPORTS = [
    "swspec_hip_derive", "swspec_hip_cfg_from_ini", "swspec_hip_create", "swspec_hip_create_multichannel",
    "swspec_hip_run", "swspec_hip_run_device", "swspec_hip_join", "swspec_hip_combine", "swspec_hip_reset",
    "swspec_hip_combine_device", "swspec_hip_reset_device",
    "swspec_hip_finalize", "swspec_hip_last_error", "swspec_hip_pcal_tones", "swspec_hip_pcal_ntones",
    "swspec_hip_last_run_seconds", "swspec_hip_host_alloc", "swspec_hip_host_free",
    "swspec_hip_unpack_device", "swspec_hip_unpack_device_noalloc", "swspec_hip_plan_info", "swspec_hip_autotune", "swspec_hip_autotune_clear", "swspec_hip_set_profiling", "swspec_hip_kernel_stats",
    "swspec_hip_server_create", "swspec_hip_server_submit", "swspec_hip_server_step", "swspec_hip_server_poll",
    "swspec_hip_server_sync", "swspec_hip_server_run_batch", "swspec_hip_server_destroy",
    "swspec_hip_server_last_error", "swspec_hip_server_stats",
    "swspec_hip_pcal_create", "swspec_hip_pcal_destroy", "swspec_hip_pcal_info", "swspec_hip_pcal_clear",
    "swspec_hip_pcal_adjust_offset", "swspec_hip_pcal_extract", "swspec_hip_pcal_final",
    "swspec_hip_pcal_copied_tones", "swspec_hip_pcal_seconds", "swspec_hip_pcal_delay",
    "swspec_hip_pcal_last_error",
]


class Job(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("window", C.c_int32), ("fft_points", C.c_int64),
        ("source_format", C.c_int32), ("source_channels", C.c_int32), ("bits_per_sample", C.c_int32),
        ("channelorder_increasing", C.c_int32), ("channel", C.c_int32), ("nffts", C.c_int32),
        ("extract_pcal", C.c_int32), ("flags", C.c_int32), ("samplingfreq", C.c_double),
        ("pcaloffsethz", C.c_double), ("pcalharmonicshz", C.c_double), ("d_raw", C.c_void_p),
        ("raw_len", C.c_uint64), ("d_out_auto", C.c_void_p), ("d_out_pcal", C.c_void_p),
        ("d_tones", C.c_void_p),
    ]


class Cfg(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("abi_version", C.c_uint32), ("fft_points", C.c_int64),
        ("fft_overlap_factor", C.c_int32), ("fft_ssb_points", C.c_int32), ("window", C.c_int32),
        ("bits_per_sample", C.c_int32), ("source_channels", C.c_int32),
        ("channelorder_increasing", C.c_int32), ("source_format", C.c_int32),
        ("use_channel_file1", C.c_int32), ("use_channel_file2", C.c_int32), ("num_sources", C.c_int32),
        ("num_xpols", C.c_int32), ("raw_fullfft_bytes", C.c_int64), ("raw_overlap_bytes", C.c_int64),
        ("core_overlapped_ffts", C.c_int32), ("max_spectra_per_buffer", C.c_int32),
        ("extract_pcal", C.c_int32), ("pcal_tonebins", C.c_int32), ("pcal_rotatorlen", C.c_int32),
        ("device", C.c_int32), ("samplingfreq", C.c_double), ("pcaloffsethz", C.c_double),
        ("pcalharmonicshz", C.c_double), ("max_rawbuf_bytes", C.c_int64), ("flags", C.c_uint32),
        ("ffts_per_chunk", C.c_uint32), ("mark5_format", C.c_char * 32),
    ]


class Ini(C.Structure):
    _fields_ = [
        ("num_cores", C.c_int32), ("max_source_buffer_mb", C.c_double), ("platform_max_raw_mb", C.c_int32),
        ("fft_points", C.c_int64), ("fft_integ_seconds", C.c_float), ("fft_overlap_factor", C.c_int32),
        ("window", C.c_int32), ("bandwidth_hz", C.c_float), ("pcal_offset_hz", C.c_float),
        ("pcal_harmonics_hz", C.c_float), ("source_format", C.c_int32), ("bits_per_sample", C.c_int32),
        ("channelorder_increasing", C.c_int32), ("source_channels", C.c_int32),
        ("use_channel_file1", C.c_int32), ("use_channel_file2", C.c_int32), ("extract_pcal", C.c_int32),
        ("num_sources", C.c_int32), ("do_xpol", C.c_int32),
    ]


class Derived(C.Structure):
    _fields_ = [
        ("averaged_ffts", C.c_int32), ("fft_overlap_points", C.c_int32), ("fft_ssb_points", C.c_int32),
        ("raw_fullfft_bytes", C.c_int64), ("raw_overlap_bytes", C.c_int64), ("rawbuf_size", C.c_int64),
        ("max_spectra_per_buffer", C.c_int32), ("max_specffts_per_buffer", C.c_int32),
        ("max_buffers_per_spectrum", C.c_int32), ("core_overlapped_ffts", C.c_int32),
        ("core_averaged_ffts", C.c_int32), ("extract_pcal", C.c_int32), ("pcal_tonebins", C.c_int32),
        ("pcal_rotatorlen", C.c_int32), ("pcal_pulses", C.c_int32), ("pcal_pulses_per_fft", C.c_int32),
        ("num_xpols", C.c_int32), ("use_channel_file1", C.c_int32), ("use_channel_file2", C.c_int32),
        ("samplingfreq", C.c_double), ("max_rawbuf_size", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load():
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libswspec_hip.so is not built; run swspec_amd.build.build() "
                           "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    PP = C.POINTER(C.c_void_p)
    SZ = C.POINTER(C.c_size_t)
    L.swspec_hip_derive.argtypes = [C.POINTER(Ini), C.POINTER(Derived)]
    L.swspec_hip_cfg_from_ini.argtypes = [C.POINTER(Ini), C.POINTER(Derived), C.POINTER(Cfg)]
    L.swspec_hip_create.argtypes = [C.POINTER(Cfg), C.c_int, C.POINTER(P)]
    L.swspec_hip_create_multichannel.argtypes = [C.POINTER(Cfg), C.c_int, C.POINTER(C.c_int), C.c_int,
                                                 C.POINTER(P)]
    L.swspec_hip_run.argtypes = [P, PP, SZ, PP, PP, PP]
    L.swspec_hip_run_device.argtypes = [P, PP, SZ, PP, PP, PP]
    L.swspec_hip_join.argtypes = [P, C.POINTER(C.c_int), SZ, SZ, SZ]
    L.swspec_hip_combine.argtypes = [P, PP, PP, SZ, SZ]
    L.swspec_hip_reset.argtypes = [P, C.c_size_t]
    L.swspec_hip_reset.restype = None
    L.swspec_hip_combine_device.argtypes = [P, PP, PP, SZ, SZ]
    L.swspec_hip_reset_device.argtypes = [P, C.c_size_t]
    L.swspec_hip_finalize.argtypes = [P]
    L.swspec_hip_last_error.argtypes = [P]
    L.swspec_hip_last_error.restype = C.c_char_p
    L.swspec_hip_pcal_tones.argtypes = [P, C.c_int, C.c_int, P, C.c_int]
    L.swspec_hip_pcal_ntones.argtypes = [P]
    L.swspec_hip_last_run_seconds.argtypes = [P]
    L.swspec_hip_last_run_seconds.restype = C.c_double
    L.swspec_hip_host_alloc.argtypes = [C.c_size_t]
    L.swspec_hip_host_alloc.restype = P
    L.swspec_hip_host_free.argtypes = [P]
    L.swspec_hip_host_free.restype = None
    L.swspec_hip_unpack_device.argtypes = [C.POINTER(Cfg), P, C.c_int, C.c_int64, P]
    L.swspec_hip_unpack_device_noalloc.argtypes = [C.POINTER(Cfg), P, C.c_int, C.c_int64, P, P]
    L.swspec_hip_plan_info.argtypes = [C.c_int64, C.c_uint32, C.POINTER(C.c_int32), C.c_int]
    L.swspec_hip_autotune.argtypes = [C.c_int, C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    L.swspec_hip_autotune_clear.restype = None
    L.swspec_hip_set_profiling.argtypes = [P, C.c_int]
    L.swspec_hip_kernel_stats.argtypes = [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.swspec_hip_server_create.argtypes = [C.c_int, C.c_int64, C.c_int, C.POINTER(P)]
    L.swspec_hip_server_submit.argtypes = [P, C.POINTER(Job), C.POINTER(C.c_int64)]
    L.swspec_hip_server_step.argtypes = [P, C.POINTER(C.c_int)]
    L.swspec_hip_server_poll.argtypes = [P, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_int)]
    L.swspec_hip_server_sync.argtypes = [P]
    L.swspec_hip_server_run_batch.argtypes = [P, C.c_int, C.c_int]
    L.swspec_hip_server_destroy.argtypes = [P]
    L.swspec_hip_server_last_error.argtypes = [P]
    L.swspec_hip_server_last_error.restype = C.c_char_p
    L.swspec_hip_server_stats.argtypes = [P, C.POINTER(C.c_int64), C.c_int]
    L.swspec_hip_pcal_create.argtypes = [C.c_double, C.c_double, C.c_int, C.c_size_t, C.c_int, C.c_int,
                                         C.POINTER(P)]
    L.swspec_hip_pcal_destroy.argtypes = [P]
    L.swspec_hip_pcal_destroy.restype = None
    L.swspec_hip_pcal_info.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int64)]
    L.swspec_hip_pcal_clear.argtypes = [P]
    L.swspec_hip_pcal_adjust_offset.argtypes = [P, C.c_size_t]
    L.swspec_hip_pcal_extract.argtypes = [P, P, C.c_size_t, P]
    L.swspec_hip_pcal_final.argtypes = [P, P, P, C.POINTER(C.c_uint64)]
    L.swspec_hip_pcal_copied_tones.argtypes = [P]
    L.swspec_hip_pcal_seconds.argtypes = [P]
    L.swspec_hip_pcal_seconds.restype = C.c_double
    L.swspec_hip_pcal_delay.argtypes = [P, C.c_int, C.POINTER(C.c_double)]
    L.swspec_hip_pcal_last_error.argtypes = [P]
    L.swspec_hip_pcal_last_error.restype = C.c_char_p
    _lib = L
    return L


def last_error(core=None):
    s = load().swspec_hip_last_error(core)
    return s.decode() if s else ""


def check(rc, core=None, what=""):
    if rc != 0:
        raise RuntimeError("%s failed (rc=%d): %s" % (what, rc, last_error(core)))
    return rc
