"""GPU: the autotuner measures the planner's distinct plans for an FFT length,
keeps the fastest, and every later plan of that length uses it -- with the
same parity bar as the planner's own plan (normwise relative error <= 1e-5 vs
the FP64 oracle)."""
import numpy as np
import pytest

from gpu_util import oracle_settings, relerr, run_device, say
from swspec_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nf,win", [(65536, 16, "Hann"), (3200000, 4, "Cosine2"), (20000, 32, "Hamming"),
                                      (1 << 20, 8, "Hann")])
def test_autotune_choice_is_measured_and_exact(hip, oracle, n, nf, win):
    before = hip.plan_info(n)
    try:
        r = hip.autotune(n, window=win, nfft=nf)
        tuned = hip.plan_info(n)
        say("autotune %d: %d plans, chosen %.3f ms vs planner %.3f ms, knobs %s; plan %s -> %s" % (
            n, r["candidates"], r["ms_best"], r["ms_default"],
            {k: r[k] for k in ("p1_maxpts", "p1_bmax", "cx_mask", "p2_maxpts")},
            [before[k] for k in ("N1", "B", "p1_threads", "p1_maxpts", "p2_maxpts", "p1_stages", "p2_stages")],
            [tuned[k] for k in ("N1", "B", "p1_threads", "p1_maxpts", "p2_maxpts", "p1_stages", "p2_stages")]))
        assert r["candidates"] >= 1 and r["ms_best"] <= r["ms_default"]
        assert tuned["N1"] * tuned["N2"] == n // 2
        fs = 32e6
        s = hip.derive_settings(FFTpoints=n, FFTIntegrationTimeSec=n * nf / fs, FFToverlapFactor=1, WindowType=win,
                                BandwidthHz=fs / 2, BitsPerSample=2, SourceChannels=1, MaxSourceBufferMB=4096)
        raw = synth.pack_1ch(synth.quantize(np.random.default_rng(n % 977).standard_normal(n * nf)))
        m, auto, _, _, core = run_device(s, [raw])
        core.finalize()
        _, a64, _, _ = oracle.domaths(oracle_settings(oracle, s), [raw], 64)
        e = relerr(auto[0], a64[0])
        say("  tuned plan vs fp64 %.3e" % e)
        assert m == 1 and e <= 1e-5
    finally:
        hip.autotune_clear()
    assert hip.plan_info(n) == before
