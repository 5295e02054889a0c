"""Build recipe of the compiled ABI check (test infrastructure): g++ compiles
tests/capi/svlg_abi_check.cpp against include/svlg.h and links libsvlg.so and the oracle's
libsvl_oracle.so, with run paths relative to the binary so the tree can move."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
SRC = HERE / "svlg_abi_check.cpp"
OUT = HERE / "svlg_abi_check"


def build() -> Path:
    import sys
    if str(ROOT) not in sys.path:
        sys.path.insert(0, str(ROOT))
    import oracle
    oracle.build()
    lib = ROOT / "svl_amd" / "libsvlg.so"
    if not lib.exists():
        raise RuntimeError(f"{lib} is not built; run `python -m svl_amd.build` first")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", str(ROOT / "include"), str(SRC), "-o", str(OUT) + ".tmp",
           "-L", str(lib.parent), "-l:libsvlg.so", "-L", str(ROOT / "oracle"), "-l:libsvl_oracle.so",
           "-Wl,-rpath,$ORIGIN/../../svl_amd", "-Wl,-rpath,$ORIGIN/../../oracle"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ABI check build failed:\n{r.stderr}")
    Path(str(OUT) + ".tmp").replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build())
