"""swspec_hip_unpack_device_noalloc -- the decoder without device
allocation (VLBA scratch in a fixed per-GPU array) -- against
swspec_hip_unpack_device through the C ABI: bit-identical samples on two
calls in a row, the second with a smaller input whose blank zones differ (the
array keeps the first call's table), for VLBA1_1-128-4-2, VLBA1_4-256-8-1 and
Mark5B-1024-8-2 (scratch used) and 4-channel Mk5B (none); a VLBA call longer than the array holds is refused
before any device work.  Bar: bit-exact."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu
PAT = np.frombuffer(bytes([0x44, 0x33, 0x22, 0x11]), dtype=np.uint8)


def _vlba_bytes(seed, n, fills):
    raw = np.random.default_rng(seed).integers(0, 256, n, dtype=np.uint8)
    raw[::4] = np.where(raw[::4] == 0x44, 0x45, raw[::4])
    for a, b in fills:
        for o in range(a, b, 4):
            raw[o:o + 4] = PAT
    return raw


def _noalloc(hip, s, raw, ch, count):
    import torch
    from swspec_amd._lib import load
    cfg = s.cfg()
    d_raw = torch.from_numpy(raw).cuda()
    ref = torch.empty(count, dtype=torch.float32, device="cuda")
    hip.unpack_device(s, d_raw, ch, count, ref)
    out = torch.full((count,), float("nan"), dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream()
    rc = load().swspec_hip_unpack_device_noalloc(C.byref(cfg), d_raw.data_ptr(), ch, count, out.data_ptr(),
                                                 C.c_void_p(st.cuda_stream))
    assert rc == 0, load().swspec_hip_last_error(None)
    st.synchronize()
    return out.cpu().numpy(), ref.cpu().numpy()


def test_noalloc_decode_matches_allocating_decode(hip):
    s = hip.derive_settings(FFTpoints=40000, FFTIntegrationTimeSec=0.01, FFToverlapFactor=1, BandwidthHz=8e6,
                            BitsPerSample=2, SourceChannels=4, SourceFormat="VLBA1_1-128-4-2")
    calls = [(4 * 40128 + 5000, [(12000, 20000), (40000 + 40128, 2 * 40128 + 8), (3 * 40128 + 20056, 3 * 40128 + 40000)]),
             (2 * 40128 + 100, [(0, 4096), (40128 + 30000, 40128 + 40000)])]
    for k, (count, fills) in enumerate(calls):
        raw = _vlba_bytes(50 + k, count, fills)
        for ch in range(4):
            got, ref = _noalloc(hip, s, raw, ch, count)
            say("vlba call %d ch %d: %d samples, %d blanked" % (k, ch, count, int(np.count_nonzero(ref == 0))))
            assert np.count_nonzero(ref == 0) > 0
            assert np.array_equal(got, ref), (k, ch, int(np.argmax(got != ref)))
    # 32 tracks (three blank zones per payload), the second call smaller
    s4 = hip.derive_settings(FFTpoints=40000, FFTIntegrationTimeSec=0.01, FFToverlapFactor=1, BandwidthHz=8e6,
                             BitsPerSample=1, SourceChannels=8, SourceFormat="VLBA1_4-256-8-1")
    for k, (count, fills) in enumerate([(2 * 160512 + 4000, [(70000, 80000), (160512 + 2000, 160512 + 40000)]),
                                        (160512 + 400, [(0, 36000)])]):
        raw = _vlba_bytes(80 + k, count, fills)
        for ch in (0, 5):
            got, ref = _noalloc(hip, s4, raw, ch, count)
            assert np.count_nonzero(ref == 0) > 0
            assert np.array_equal(got, ref), (k, ch)
    # Mark5B bitstreams (one blank zone per 10000-byte payload)
    s5 = hip.derive_settings(FFTpoints=40000, FFTIntegrationTimeSec=0.01, FFToverlapFactor=1, BandwidthHz=8e6,
                             BitsPerSample=2, SourceChannels=8, SourceFormat="Mark5B-1024-8-2")
    for k, (count, fills) in enumerate([(20000, [(25000, 30000)]), (6000, [(0, 4096)])]):
        raw = _vlba_bytes(90 + k, 2 * count, fills)
        got, ref = _noalloc(hip, s5, raw, 3, count)
        assert np.count_nonzero(ref == 0) > 0
        assert np.array_equal(got, ref), k
    m = hip.derive_settings(FFTpoints=4000, FFTIntegrationTimeSec=0.01, FFToverlapFactor=1, BandwidthHz=8e6,
                            BitsPerSample=2, SourceChannels=4, SourceFormat="Mk5B")
    for k, count in enumerate((60000, 20000)):
        raw = np.random.default_rng(60 + k).integers(0, 256, count, dtype=np.uint8)
        got, ref = _noalloc(hip, m, raw, 2, count)
        assert np.array_equal(got, ref), k


def test_noalloc_refuses_vlba_beyond_scratch(hip):
    import torch
    from swspec_amd._lib import load
    s = hip.derive_settings(FFTpoints=40000, FFTIntegrationTimeSec=0.01, FFToverlapFactor=1, BandwidthHz=8e6,
                            BitsPerSample=2, SourceChannels=4, SourceFormat="VLBA1_1-128-4-2")
    cfg = s.cfg()
    d = torch.zeros(16, dtype=torch.uint8, device="cuda")
    o = torch.zeros(16, dtype=torch.float32, device="cuda")
    rc = load().swspec_hip_unpack_device_noalloc(C.byref(cfg), d.data_ptr(), 0, 43584 * 40128 + 1, o.data_ptr(),
                                                 None)
    assert rc == -1
    assert b"scratch" in load().swspec_hip_last_error(None)
