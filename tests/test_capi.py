"""CPU checks of the C ABI: the library loads and exports every declared entry point."""
import re
from pathlib import Path

from svl_amd import _capi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "svlg.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(svlg_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_table():
    assert declared_functions() == sorted(_capi.SIGNATURES)


def test_library_exports_every_symbol():
    L = _capi.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_version_and_device_count_without_gpu():
    L = _capi.lib()
    assert b"gfx950" in L.svlg_version()
    assert _capi.device_count() >= 0


def test_cpp_caller_compiles_against_the_header_and_links():
    """A C++ program (tests/capi/svlg_abi_check.cpp) compiles with g++ against include/svlg.h
    alone and links libsvlg.so: every entry point it calls is declared in the header and exported
    by the library (the GPU test runs it, which checks what the calls compute)."""
    import subprocess
    import sys
    sys.path.insert(0, str(ROOT / "tests" / "capi"))
    import build_abi_check
    out = build_abi_check.build()
    assert out.exists()
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(out)], capture_output=True, text=True).stdout
    used = sorted(set(w for w in nm.split() if w.startswith("svlg_")))
    assert len(used) >= 15 and set(used) <= set(declared_functions())
