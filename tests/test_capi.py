"""The C-ABI library loads on the CPU host and exports every declared symbol.
No compute calls here (no GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(swspec_hip_\w+)\s*\(", text):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_all_declared_symbols():
    from swspec_amd import _lib
    L = _lib.load()
    decl = declared_functions()
    assert len(decl) >= 18
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == decl


def test_cfg_struct_layout_matches_header():
    # sizeof(swspec_hip_cfg) as the C compiler lays it out
    from swspec_amd import _lib
    assert C.sizeof(_lib.Cfg) == 176
    assert C.sizeof(_lib.Ini) == 96
    assert C.sizeof(_lib.Derived) == 112


def test_derive_errors_and_plan_info():
    import pytest
    import swspec_amd
    with pytest.raises(RuntimeError):
        swspec_amd.derive_settings(UseFile1Channel=3, SourceChannels=2, BitsPerSample=2)
    with pytest.raises(RuntimeError):
        swspec_amd.derive_settings(ExtractPCal=True, PCalOffsetHz=0.0)
    info = swspec_amd.plan_info(3200000)
    assert info["N1"] * info["N2"] == 1600000
    # rows 5*16*16: the radix-16 stage after the radix-5 one (ns = 5, blocks of
    # 80 points) writes padded blocks; 2^20 rows 4*16*16 likewise (ns = 4);
    # 32000 has two row stages and no padding
    assert info["p2_pad"] == 5
    assert swspec_amd.plan_info(1 << 20)["p2_pad"] == 4
    assert swspec_amd.plan_info(32000)["p2_pad"] == 0
    with pytest.raises(RuntimeError):
        swspec_amd.plan_info(2 * 7 * 1000)   # prime factor 7: no plan
    assert info["mfma64"] == 0


@pytest.mark.parametrize("n", [1 << 20, 3200000, 65536, 32000, 8192, 4096, 320000, 1 << 22])
def test_plan_info_mfma64(n):
    """Radix-64 MFMA stage geometry: a column length N1 that is a multiple
    of 64 and divides N/2, stages [64, ...] multiplying to N1, one pipeline
    of whole waves with at most 8 (row tile, column tile) pairs per wave."""
    import swspec_amd
    from swspec_amd._lib import FLAG_MFMA64
    i = swspec_amd.plan_info(n, FLAG_MFMA64)
    assert i["mfma64"] == 1 and i["p1_radix0"] == 64
    assert i["N1"] % 64 == 0 and i["N1"] * i["N2"] == n // 2 and i["N2"] % i["B"] == 0
    assert i["p1_threads"] % 64 == 0 and i["p1_threads"] <= 512 and i["p1_maxpts"] == 16
    waves = i["p1_threads"] // 64
    tiles = ((i["N1"] // 64) * i["B"] + 15) // 16
    assert -(-8 // waves) * tiles <= 8
    assert i["N1"] * i["B"] * 8 <= 80 * 1024


@pytest.mark.parametrize("n", [20000, 1000000])
def test_plan_info_mfma64_rejects(n):
    """N/2 = 10^4 and 5 x 10^5 have no divisor that is a multiple of 64."""
    import swspec_amd
    from swspec_amd._lib import FLAG_MFMA64
    with pytest.raises(RuntimeError):
        swspec_amd.plan_info(n, FLAG_MFMA64)
