"""The phase-cal extractor classes (src/IA-32/PhaseCal/PCal.h:51-190) on the GPU.

PCal.getNew(bandwidth_hz, pcal_spacing_hz, pcal_offset_hz, sampleoffset)
returns the class the reference's factory picks (PCal.cpp:91-110):
PCalExtractorTrivial for a zero offset, PCalExtractorImplicitShift when the
offset's repeat length is a multiple of the comb's, else
PCalExtractorShifting.  Methods keep the reference's names and meaning:
clear, adjustSampleOffset, extractAndIntegrate, getSeconds, getLength,
getNBins, getFinalPCal.  Samples are float32; a CUDA tensor is read in place
on torch's current stream, anything else is copied to the device first.
Accumulation is FP64 on the device (swspec_hip_pcal_* in
include/swspec_hip.h); tones come back as complex64 (getFinalPCal) or
complex128 (getFinalPCal64).  delay_seconds() adds the time-domain delay
analysis that PCal.h:15-20 describes.
"""
import ctypes as C

import numpy as np

from ._lib import load

KINDS = {0: "Trivial", 1: "Shifting", 2: "ImplicitShift"}
_KIND_IDS = {v: k for k, v in KINDS.items()}


def _err(p=None):
    s = load().swspec_hip_pcal_last_error(p)
    return s.decode() if s else ""


def _check(rc, p, what):
    if rc != 0:
        raise RuntimeError("%s failed (rc=%d): %s" % (what, rc, _err(p)))


class PCal:
    """One extractor object (PCal.h); use PCal.getNew or PCal.of_kind."""

    def __init__(self, bandwidth_hz, pcal_spacing_hz, pcal_offset_hz, sampleoffset=0, kind=None, device=0):
        import torch
        torch.cuda.init()   # torch's HIP runtime first: it then shares the device with this library
        h = C.c_void_p()
        k = -1 if kind is None else _KIND_IDS[kind]
        rc = load().swspec_hip_pcal_create(float(bandwidth_hz), float(pcal_spacing_hz), int(pcal_offset_hz),
                                           int(sampleoffset), k, int(device), C.byref(h))
        _check(rc, None, "PCal create")
        self._p = h
        self.device = device
        kd, nt, nb, L = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        _check(load().swspec_hip_pcal_info(h, C.byref(kd), C.byref(nt), C.byref(nb), C.byref(L)), h, "PCal info")
        self.kind = KINDS[kd.value]
        self._ntones, self._nbins, self.rotatorlen = nt.value, nb.value, L.value
        self.spacing_hz = int(pcal_spacing_hz) if self.kind == "Trivial" else float(pcal_spacing_hz)

    @staticmethod
    def getNew(bandwidth_hz, pcal_spacing_hz, pcal_offset_hz, sampleoffset=0, device=0):
        """PCal::getNew (PCal.cpp:91-110)."""
        return PCal(bandwidth_hz, pcal_spacing_hz, pcal_offset_hz, sampleoffset, None, device)

    @staticmethod
    def of_kind(kind, bandwidth_hz, pcal_spacing_hz, pcal_offset_hz=0, sampleoffset=0, device=0):
        """Direct construction of one class ("Trivial", "Shifting", "ImplicitShift")."""
        return PCal(bandwidth_hz, pcal_spacing_hz, pcal_offset_hz, sampleoffset, kind, device)

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            load().swspec_hip_pcal_destroy(p)
            self._p = None

    def clear(self):
        _check(load().swspec_hip_pcal_clear(self._p), self._p, "clear")

    def adjustSampleOffset(self, sampleoffset):
        _check(load().swspec_hip_pcal_adjust_offset(self._p, int(sampleoffset)), self._p, "adjustSampleOffset")

    def extractAndIntegrate(self, samples, length=None):
        """False once finalized (as the reference); raises on other errors."""
        import torch
        if isinstance(samples, torch.Tensor) and samples.is_cuda:
            t = samples
        else:
            t = torch.as_tensor(np.ascontiguousarray(samples, dtype=np.float32)).to(f"cuda:{self.device}")
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("samples must be contiguous float32")
        n = t.numel() if length is None else int(length)
        if n > t.numel():
            raise ValueError("length exceeds the samples given")
        st = torch.cuda.current_stream(t.device).cuda_stream
        rc = load().swspec_hip_pcal_extract(self._p, t.data_ptr(), n, st)
        if rc == -5:   # SWSPEC_ESTATE: finalized
            return False
        _check(rc, self._p, "extractAndIntegrate")
        return True

    def getSeconds(self):
        return load().swspec_hip_pcal_seconds(self._p)

    def getLength(self):
        return self._ntones

    def getNBins(self):
        return self._nbins

    def copied_tones(self):
        """Tones getFinalPCal writes (the reference stops at the last bin)."""
        return load().swspec_hip_pcal_copied_tones(self._p)

    def getFinalPCal(self, out=None):
        """Returns (samplecount, tones complex64).  `out` (complex64, getLength
        entries) is written in place; entries past copied_tones() keep their
        values, as in the reference."""
        if out is None:
            out = np.zeros(self._ntones, dtype=np.complex64)
        if out.dtype != np.complex64 or len(out) < self._ntones or not out.flags.c_contiguous:
            raise ValueError("out must be contiguous complex64 of getLength() entries")
        n = C.c_uint64()
        _check(load().swspec_hip_pcal_final(self._p, out.ctypes.data, None, C.byref(n)), self._p, "getFinalPCal")
        return int(n.value), out

    def getFinalPCal64(self):
        """FP64 tones (the first copied_tones() entries are defined)."""
        out = np.zeros(self._ntones, dtype=np.complex128)
        n = C.c_uint64()
        _check(load().swspec_hip_pcal_final(self._p, None, out.ctypes.data, C.byref(n)), self._p, "getFinalPCal")
        return int(n.value), out

    def delay_seconds(self, upsample=64):
        """Delay of the comb within one period 1/spacing, from the peak of its
        time-domain response (needs getFinalPCal first)."""
        d = C.c_double()
        _check(load().swspec_hip_pcal_delay(self._p, int(upsample), C.byref(d)), self._p, "delay")
        return d.value
