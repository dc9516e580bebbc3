"""Batched and served execution (swspec_hip_server_* of include/swspec_hip.h).

No counterpart in the reference, which runs one job per process.  A job is one
channel of a device-resident stream integrated over `nffts` FFTs into one
spectrum (scale 1/nffts), optionally with phase-cal.
"""
import ctypes as C

from ._lib import FORMATS, WINDOWS, Job, load

EAGAIN = -6


class ServerError(RuntimeError):
    pass


class Server:
    def __init__(self, device=0, hbm_budget_bytes=0, max_active=256):
        self._s = C.c_void_p()
        rc = load().swspec_hip_server_create(device, hbm_budget_bytes, max_active, C.byref(self._s))
        if rc != 0:
            raise ServerError("swspec_hip_server_create failed (rc=%d)" % rc)
        self._keep = {}

    def _err(self, what, rc):
        msg = load().swspec_hip_server_last_error(self._s)
        return ServerError("%s failed (rc=%d): %s" % (what, rc, msg.decode() if msg else ""))

    def submit(self, fft_points, nffts, d_raw, raw_len, d_out_auto, window="Cosine2", source_format="RawSigned",
               source_channels=1, channel=0, samplingfreq=32e6, extract_pcal=False, pcaloffsethz=10e3,
               pcalharmonicshz=1e6, d_out_pcal=None, d_tones=None, channelorder_increasing=True, keep=None,
               flags=0):
        """Admit a job; returns its id, or None when the server is full (retry later)."""
        def ptr(x):
            return None if x is None else (x if isinstance(x, int) else x.data_ptr())
        j = Job()
        j.struct_size = C.sizeof(Job)
        j.window = WINDOWS[window] if isinstance(window, str) else int(window)
        j.fft_points = int(fft_points)
        j.source_format = FORMATS[source_format] if isinstance(source_format, str) else int(source_format)
        j.source_channels = int(source_channels)
        j.bits_per_sample = 2
        j.channelorder_increasing = 1 if channelorder_increasing else 0
        j.channel = int(channel)
        j.nffts = int(nffts)
        j.extract_pcal = 1 if extract_pcal else 0
        j.flags = int(flags)
        j.samplingfreq = float(samplingfreq)
        j.pcaloffsethz = float(pcaloffsethz)
        j.pcalharmonicshz = float(pcalharmonicshz)
        j.d_raw = ptr(d_raw)
        j.raw_len = int(raw_len)
        j.d_out_auto = ptr(d_out_auto)
        j.d_out_pcal = ptr(d_out_pcal)
        j.d_tones = ptr(d_tones)
        jid = C.c_int64(0)
        rc = load().swspec_hip_server_submit(self._s, C.byref(j), C.byref(jid))
        if rc == EAGAIN:
            return None
        if rc != 0:
            raise self._err("swspec_hip_server_submit", rc)
        if keep is not None:
            self._keep[jid.value] = keep
        return jid.value

    def step(self):
        n = C.c_int(0)
        rc = load().swspec_hip_server_step(self._s, C.byref(n))
        if rc != 0:
            raise self._err("swspec_hip_server_step", rc)
        return n.value

    def poll(self, max_ids=4096):
        ids = (C.c_int64 * max_ids)()
        n = C.c_int(0)
        rc = load().swspec_hip_server_poll(self._s, ids, max_ids, C.byref(n))
        if rc != 0:
            raise self._err("swspec_hip_server_poll", rc)
        out = list(ids[: n.value])
        for i in out:
            self._keep.pop(i, None)
        return out

    def sync(self):
        rc = load().swspec_hip_server_sync(self._s)
        if rc != 0:
            raise self._err("swspec_hip_server_sync", rc)

    def run_batch(self, use_graph=True, replays=1):
        rc = load().swspec_hip_server_run_batch(self._s, 1 if use_graph else 0, replays)
        if rc != 0:
            raise self._err("swspec_hip_server_run_batch", rc)

    STATS = ("steps", "cross_stream_waits", "admitted", "completed", "active", "resident_bytes")

    def stats(self):
        """Counters of swspec_hip_server_stats as a dict (names in STATS)."""
        v = (C.c_int64 * len(self.STATS))()
        rc = load().swspec_hip_server_stats(self._s, v, len(self.STATS))
        if rc != 0:
            raise self._err("swspec_hip_server_stats", rc)
        return dict(zip(self.STATS, list(v)))

    def close(self):
        if self._s:
            load().swspec_hip_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
