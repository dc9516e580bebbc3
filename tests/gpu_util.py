"""Helpers for the GPU parity tests: run TaskCoreHIP on device-resident inputs."""
import sys

import numpy as np


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nspectra_for(settings, nbytes):
    """doMaths loop arithmetic (TaskCoreIPP.cpp:394-524) -> number of spectra."""
    full, ovl, K = settings.raw_fullfft_bytes, settings.raw_overlap_bytes, settings.core_overlapped_ffts
    rem, acc, sp = nbytes, 0, 0
    while rem >= full:
        rem -= ovl
        acc += 1
        if acc >= K:
            sp += 1
            acc = 0
            rem -= full - ovl
    return sp + (1 if acc else 0)


def run_device(settings, raws, channels=None, flags=0, ffts_per_chunk=0):
    """Returns (nspectra, [auto per job], [xpol] or [], [pcal per job], core)."""
    import torch
    from swspec_amd import TaskCoreHIP
    core = TaskCoreHIP(0)
    core.prepare(settings, flags=flags, ffts_per_chunk=ffts_per_chunk, channels=channels)
    dev = torch.device("cuda:0")
    d_in = [torch.from_numpy(np.ascontiguousarray(r)).to(dev) for r in raws]
    nbytes = min(len(r) for r in raws)
    nsp = max(nspectra_for(settings, nbytes), 1)
    njobs = core.njobs
    ssb = settings.fft_ssb_points
    d_auto = [torch.full((nsp * ssb,), float("nan"), dtype=torch.float32, device=dev) for _ in range(njobs)]
    d_x = [torch.full((2 * nsp * ssb,), float("nan"), dtype=torch.float32, device=dev)] if settings.num_xpols else None
    T = settings.pcal_tonebins if settings.extract_pcal else 0
    d_p = [torch.full((2 * nsp * max(T, 1),), float("nan"), dtype=torch.float32, device=dev)
           for _ in range(njobs)] if T else None
    core.run_device(d_in, [len(r) for r in raws], d_auto, d_x, d_p)
    n = core.join()
    assert n >= 0
    auto = [a.cpu().numpy().reshape(nsp, ssb)[:n] for a in d_auto]
    xp = []
    if d_x:
        v = d_x[0].cpu().numpy()
        xp = [(v[0::2] + 1j * v[1::2]).reshape(nsp, ssb)[:n]]
    pc = []
    if d_p:
        for p in d_p:
            v = p.cpu().numpy()
            pc.append((v[0::2] + 1j * v[1::2]).reshape(nsp, T)[:n])
    return n, auto, xp, pc, core


def relerr(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def oracle_settings(oracle, s):
    """Oracle settings for the same INI keys as a product Settings object."""
    k = s.keys
    return oracle.derive(
        num_cores=int(k["NumCPUCores"]),
        max_source_buffer_mb=-1.0 if k["MaxSourceBufferMB"] is None else float(k["MaxSourceBufferMB"]),
        fft_points=int(k["FFTpoints"]), fft_integ_seconds=float(k["FFTIntegrationTimeSec"]),
        fft_overlap_factor=int(k["FFToverlapFactor"]), window=k["WindowType"],
        bandwidth_hz=float(k["BandwidthHz"]), pcal_offset_hz=float(k["PCalOffsetHz"]),
        source_format=k["SourceFormat"], bits_per_sample=int(k["BitsPerSample"]),
        channelorder_increasing=1 if k["ChannelOrderIncreasing"] else 0,
        source_channels=int(k["SourceChannels"]), use_channel_file1=int(k["UseFile1Channel"]),
        use_channel_file2=int(k["UseFile2Channel"]), extract_pcal=1 if k["ExtractPCal"] else 0,
        num_sources=s.num_sources, do_xpol=1 if k["DoCrossPolarization"] else 0)


def oracle_split(oracle, mk, raws, nfft, bytes_per_fft, parts=8, precision=64):
    """The oracle's doMaths over `nfft` non-overlapped FFTs (one spectrum,
    K = nfft) computed as `parts` contiguous sub-integrations in threads
    (ctypes releases the GIL while the C oracle runs) and combined by the
    combine identity of TaskDispatcher (TaskCoreIPP.cpp:124, 311-349): the
    mean is sum_p (n_p / nfft) mean_p, and phase-cal folds (raw sums) add.
    mk(n) returns oracle settings integrating n FFTs; raws are the sources'
    byte arrays.  Returns (auto per source, xpol list, pcal per source), FP64."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    parts = max(1, min(parts, nfft))
    edges = [round(nfft * i / parts) for i in range(parts + 1)]

    def one(i):
        a, b = edges[i], edges[i + 1]
        s = mk(b - a)
        assert s.core_overlapped_ffts == b - a, (s.core_overlapped_ffts, b - a)
        sub = [np.ascontiguousarray(r[a * bytes_per_fft: b * bytes_per_fft]) for r in raws]
        m, auto, xp, pc = oracle.domaths(s, sub, precision)
        assert m == 1
        return b - a, auto, xp, pc

    nth = max(1, min(parts, (os.cpu_count() or 1), 16))
    with ThreadPoolExecutor(max_workers=nth) as ex:
        res = list(ex.map(one, range(parts)))
    w = [n / nfft for n, _, _, _ in res]
    auto = [sum(wi * r[1][k][0].astype(np.float64) for wi, r in zip(w, res)) for k in range(len(res[0][1]))]
    xp = [sum(wi * r[2][k][0].astype(np.complex128) for wi, r in zip(w, res)) for k in range(len(res[0][2]))]
    pc = [sum(r[3][k][0].astype(np.complex128) for r in res) for k in range(len(res[0][3]))]
    return auto, xp, pc
