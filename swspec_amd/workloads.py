"""Synthetic multi-job workloads for the batched and served engine.

Streams are generated on the device from a seed (Gaussian noise + tones,
quantised at +-0.98 sigma with the reference's 2-bit field convention,
DataUnpackers.cpp:212) and stored as rings of a few FFTs that jobs read
cyclically (the engine takes FFT k from slot k mod ring length), or, for
throughput runs, as whole per-stream inputs of random bytes (full_inputs).

  batch_workload : 64 streams, FFT sizes drawn from {4096, 20000, 65536,
                   320000, 2^20, 3.2M}, 10-5000 FFTs each, raw 1-channel or
                   Mk5B 4-channel at 50/50, one random tone per stream.
  serving_trace  : Poisson arrivals; each job 16/32/64 Ms/s, 1-4 channels,
                   FFT size from the same set, 0.1-10 s, optional phase-cal
                   (where N is a multiple of the rotator length).
"""
import math
import sys
import time

import numpy as np

SIZES = [4096, 20000, 65536, 320000, 1 << 20, 3200000]


def _fields(torch, x, t=0.98):
    return torch.where(x >= t, 2, torch.where(x >= 0, 0, torch.where(x >= -t, 1, 3))).to(torch.uint8)


def _noise_tones(torch, g, n, s0, fs, tones, dev):
    x = torch.randn(n, generator=g, device=dev, dtype=torch.float32)
    if tones:
        idx = torch.arange(s0, s0 + n, device=dev, dtype=torch.int64)
        for (a, f, ph) in tones:
            fi = int(round(f))
            r = (idx * fi) % int(fs)
            x += a * torch.cos(r.to(torch.float32) * (2 * math.pi / fs) + ph)
    return x


def make_ring(torch, seed, fmt, nsamples, fs, tones_per_ch, dev="cuda"):
    """Device uint8 ring: fmt 'raw1' (4 samples/byte, oldest in MSB) or 'mk5b4'
    (one byte per time sample, channel c in bits 2c..2c+1)."""
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    slab = 1 << 26
    if fmt == "raw1":
        assert nsamples % 4 == 0
        out = torch.empty(nsamples // 4, dtype=torch.uint8, device=dev)
        for s0 in range(0, nsamples, slab):
            n = min(slab, nsamples - s0)
            f = _fields(torch, _noise_tones(torch, g, n, s0, fs, tones_per_ch[0], dev)).view(-1, 4)
            out[s0 // 4: (s0 + n) // 4] = (f[:, 0] << 6) | (f[:, 1] << 4) | (f[:, 2] << 2) | f[:, 3]
        return out
    out = torch.zeros(nsamples, dtype=torch.uint8, device=dev)
    for ch in range(4):
        for s0 in range(0, nsamples, slab):
            n = min(slab, nsamples - s0)
            f = _fields(torch, _noise_tones(torch, g, n, s0, fs, tones_per_ch[ch], dev))
            out[s0: s0 + n] |= f << (2 * ch)
    return out


def comb_tones(rng, nch, ntones=16, amp=0.141, offset=10_000, spacing=1_000_000):
    """Per channel: a phase-cal comb of `ntones` tones at offset + k*spacing Hz
    (amplitude 0.141 = -20 dB of unit noise power), seeded phases."""
    return [[(amp, offset + k * spacing, float(rng.uniform(0, 2 * np.pi))) for k in range(ntones)]
            for _ in range(nch)]


def mk5b_payload(torch, seed, nsamples, fs=32e6, nch=4, ntones=16, dev="cuda"):
    """Seeded 4-channel 2-bit payload on the device: one byte per time sample,
    channel c in bits 2c..2c+1 (Mk5BUnpacker's increasing order), each channel
    unit Gaussian noise plus a phase-cal comb; returns (uint8 tensor, tones)."""
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    raw = torch.zeros(nsamples, dtype=torch.uint8, device=dev)
    tones = comb_tones(np.random.default_rng(seed), nch, ntones)
    ifs = int(fs)
    slab = 40_000_000
    for ch in range(nch):
        for s0 in range(0, nsamples, slab):
            n = min(slab, nsamples - s0)
            x = torch.randn(n, generator=g, device=dev, dtype=torch.float32)
            idx = torch.arange(s0, s0 + n, device=dev, dtype=torch.int64)
            for (a, f, ph) in tones[ch]:
                r = (idx * f) % ifs          # exact phase numerator
                x += a * torch.cos(r.to(torch.float32) * (2 * math.pi / fs) + ph)
            raw[s0:s0 + n] |= _fields(torch, x) << (2 * ch)
            del x, idx
    torch.cuda.synchronize()
    return raw, tones


def xpol_pair(torch, seed, n, nfft, rho=0.2, phase_deg=30.0, dev="cuda"):
    """Cross-pol pair as two 1-channel 2-bit streams (4 samples per byte,
    oldest in the MSB): per FFT block of n samples, X ~ N(0,1) and
    Y = rho (cos(phi) X + sin(phi) H{X}) + sqrt(1 - rho^2) Z, H the Hilbert
    transform of the block (FP64 FFT), Z independent noise."""
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    ph = math.radians(phase_deg)
    out = [torch.empty(n // 4 * nfft, dtype=torch.uint8, device=dev) for _ in range(2)]
    f = torch.fft.fftfreq(n, device=dev, dtype=torch.float64)
    hmul = -1j * torch.sign(f)                  # Hilbert transform in frequency
    blk = max(1, (1 << 25) // n)
    for b0 in range(0, nfft, blk):
        nb = min(blk, nfft - b0)
        x = torch.randn(nb, n, generator=g, device=dev, dtype=torch.float64)
        z = torch.randn(nb, n, generator=g, device=dev, dtype=torch.float64)
        h = torch.fft.ifft(torch.fft.fft(x, dim=1) * hmul, dim=1).real
        y = rho * (math.cos(ph) * x + math.sin(ph) * h) + math.sqrt(1 - rho * rho) * z
        for k, v in enumerate((x, y)):
            q = _fields(torch, v.reshape(-1)).view(-1, 4)
            out[k][b0 * n // 4:(b0 + nb) * n // 4] = (q[:, 0] << 6) | (q[:, 1] << 4) | (q[:, 2] << 2) | q[:, 3]
        del x, z, h, y
    torch.cuda.synchronize()
    return out


def batch_workload(seed=4, nstreams=64, fs=32e6):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nstreams):
        n = SIZES[int(rng.integers(0, len(SIZES)))]
        out.append(dict(fft_points=n, nffts=int(rng.integers(10, 5001)),
                        fmt="raw1" if rng.random() < 0.5 else "mk5b4",
                        channel=int(rng.integers(0, 4)), tone=float(rng.uniform(0.05, 0.95) * fs / 2),
                        seed=int(rng.integers(0, 2**31))))
    return out


def raw_bytes_per_fft(fmt, n):
    return n // 4 if fmt == "raw1" else n


class RingPool:
    """Rings keyed by (fmt, N, fs, tone): ring_ffts FFTs each."""

    def __init__(self, torch, ring_ffts=8, dev="cuda"):
        self.torch, self.ring_ffts, self.dev = torch, ring_ffts, dev
        self.rings = {}

    def get(self, fmt, n, fs, tones_key, tones_per_ch, seed):
        key = (fmt, n, fs, tones_key)
        if key not in self.rings:
            self.rings[key] = make_ring(self.torch, seed, fmt, n * self.ring_ffts, fs, tones_per_ch, self.dev)
        return self.rings[key]


def full_inputs(torch, streams, seed=11, dev="cuda"):
    """Every stream's whole input (nffts FFTs, no ring) in one device buffer,
    each at a 256-byte aligned offset, filled with uniform random bytes (every
    2-bit field equally likely; the kernels' work does not depend on the
    values).  Returns (buffer, [(offset, bytes)] per stream)."""
    offs, o = [], 0
    for s in streams:
        b = s["nffts"] * raw_bytes_per_fft(s["fmt"], s["fft_points"])
        offs.append((o, b))
        o += (b + 255) // 256 * 256
    buf = torch.empty(o, dtype=torch.uint8, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    slab = 1 << 32
    for a in range(0, o, slab):
        buf[a:a + slab].random_(0, 256, generator=g)
    torch.cuda.synchronize()
    return buf, offs


def submit_batch(torch, server, streams, pool, fs=32e6, window="Cosine2", full=None):
    """Admit every stream of a batch workload; returns (outputs, total samples).
    full: (buffer, offsets) of full_inputs, read instead of the pool's rings."""
    outs, total = [], 0
    for i, s in enumerate(streams):
        if full is not None:
            ring = full[0].data_ptr() + full[1][i][0]
            nbytes = full[1][i][1]
        else:
            tones = [[(0.1, s["tone"], 0.0)]] * 4
            ring = pool.get(s["fmt"], s["fft_points"], fs, round(s["tone"]), tones, s["seed"])
            nbytes = ring.numel()
        out = torch.empty(s["fft_points"] // 2 + 1, dtype=torch.float32, device=pool.dev)
        jid = server.submit(s["fft_points"], s["nffts"], ring, nbytes, out, window=window,
                            source_format="RawSigned" if s["fmt"] == "raw1" else "Mk5B",
                            source_channels=1 if s["fmt"] == "raw1" else 4,
                            channel=0 if s["fmt"] == "raw1" else s["channel"], samplingfreq=fs)
        if jid is None:
            raise RuntimeError("batch job not admitted (server capacity)")
        outs.append(out)
        total += s["nffts"] * s["fft_points"]
    return outs, total


def serving_trace(seed=5, rate=200.0, duration=60.0):
    """Poisson arrivals: list of job dicts with arrival time t (s)."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        fs = float([16e6, 32e6, 64e6][int(rng.integers(0, 3))])
        n = SIZES[int(rng.integers(0, len(SIZES)))]
        T = float(rng.uniform(0.1, 10.0))
        nch = int(rng.integers(1, 5))
        L = int(fs / math.gcd(int(fs), 10000))
        pcal = bool(rng.random() < 0.5) and (n % L == 0)
        jobs.append(dict(t=t, fs=fs, fft_points=n, nffts=max(1, int(fs * T / n)), nch=nch, pcal=pcal))
    return jobs


def run_serving(torch, server, trace, pool, window="Cosine2", time_scale=1.0, max_wall=None,
                max_active_channels=256, max_active_jobs=None):
    """Replay a Poisson trace against the server in real time (scaled).

    Continuous batching: jobs that arrive while others run are admitted (all
    their channels at once) as soon as the active-channel cap allows, and then
    advance in the same steps as the running ones.  Returns the latencies (s)
    in completion order, total channel-samples, the wall time from the start
    to the last completion, and `jobs`: the trace's jobs, each completed one
    carrying its "latency" (arrival -> completion) and "wait" (arrival ->
    admission)."""
    pending = sorted((dict(j) for j in trace), key=lambda j: j["t"])   # per-run bookkeeping
    queue, live, lat = [], {}, []
    outs = {}
    total = 0
    active_ch = 0
    t0 = time.perf_counter()
    t_log = t0
    i = 0
    ndone = 0
    done_total = 0
    truncated = False
    while ndone < len(pending):
        if time.perf_counter() - t_log > 10.0:
            t_log = time.perf_counter()
            print("serving: t=%.1fs arrived=%d queued=%d active_ch=%d done=%d" % (
                t_log - t0, i, len(queue), active_ch, ndone), file=sys.stderr, flush=True)
        if max_wall is not None and time.perf_counter() - t0 > max_wall:
            truncated = True
            break
        now = (time.perf_counter() - t0) / time_scale
        while i < len(pending) and pending[i]["t"] <= now:
            queue.append(pending[i])
            i += 1
        # admit whole jobs (all channels) in arrival order
        while (queue and active_ch + queue[0]["nch"] <= max_active_channels
               and (max_active_jobs is None or len(live) < max_active_jobs)):
            jb = queue.pop(0)
            jb["wait"] = (time.perf_counter() - t0) / time_scale - jb["t"]   # arrival -> admission
            fmt = "raw1" if jb["nch"] == 1 else "mk5b4"
            tones = [[(0.1, 4.8e6 * jb["fs"] / 32e6, 0.0)]] * 4
            ring = pool.get(fmt, jb["fft_points"], jb["fs"], 0, tones, 17 + jb["fft_points"])
            ids = []
            for ch in range(jb["nch"]):
                o = torch.empty(jb["fft_points"] // 2 + 1, dtype=torch.float32, device=pool.dev)
                jid = server.submit(jb["fft_points"], jb["nffts"], ring, ring.numel(), o, window=window,
                                    source_format="RawSigned" if fmt == "raw1" else "Mk5B",
                                    source_channels=1 if fmt == "raw1" else 4, channel=ch, samplingfreq=jb["fs"],
                                    extract_pcal=jb["pcal"], keep=o)
                if jid is None:
                    raise RuntimeError("server refused a job inside the channel cap (memory budget)")
                ids.append(jid)
            active_ch += len(ids)
            live[tuple(ids)] = jb
            for jid in ids:
                outs[jid] = tuple(ids)
            total += jb["nffts"] * jb["fft_points"] * jb["nch"]
        nact = server.step()
        for jid in server.poll():
            key = outs.pop(jid, None)
            if key is None:
                continue
            active_ch -= 1
            jb = live.get(key)
            if jb is None:
                continue
            jb.setdefault("left", len(key))
            jb["left"] -= 1
            if jb["left"] == 0:
                jb["latency"] = (time.perf_counter() - t0) / time_scale - jb["t"]
                lat.append(jb["latency"])
                ndone += 1
                done_total += jb["nffts"] * jb["fft_points"] * jb["nch"]
                del live[key]
        if nact == 0 and not queue:
            if i < len(pending):
                dt = pending[i]["t"] - (time.perf_counter() - t0) / time_scale
                if dt > 0:
                    # with jobs still on the device, wake often enough to see
                    # them complete (a sleep to the next arrival would add up
                    # to the inter-arrival gap to their latency)
                    time.sleep(min(dt * time_scale, 2e-4 if live else 0.05))
            else:
                server.sync()
    server.sync()
    for jid in server.poll():
        key = outs.pop(jid, None)
        if key and key in live:
            jb = live[key]
            jb.setdefault("left", len(key))
            jb["left"] -= 1
            if jb["left"] == 0:
                jb["latency"] = (time.perf_counter() - t0) / time_scale - jb["t"]
                lat.append(jb["latency"])
                ndone += 1
                done_total += jb["nffts"] * jb["fft_points"] * jb["nch"]
                del live[key]
    wall = time.perf_counter() - t0
    # throughput counts completed jobs only (a truncated replay leaves some running)
    return dict(latencies=lat, samples=done_total, admitted_samples=total, wall=wall, completed=ndone,
                submitted=len(pending), truncated=truncated, jobs=pending)
