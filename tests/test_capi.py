"""The C-ABI library loads on the CPU host and exports every declared symbol.
No compute calls here (no GPU)."""
import ctypes as C
import os
import re

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
    with pytest.raises(RuntimeError):
        swspec_amd.plan_info(2 * 7 * 1000)   # prime factor 7: no plan
