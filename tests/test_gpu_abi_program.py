"""GPU: the compiled C++ caller of the ABI (tests/capi/svlg_abi_check.cpp, built by g++ against
include/svlg.h with __graft_entry__.build()) runs a Quad4 layer through the fused device steps
and through the Integrator methods one by one (ComputeEffectiveForce, ComputeEffectiveStiffness,
the solve, ComputeNewStep's update) for CentralDifference and ExtendedNewmarkBeta (lumped and
consistent mass), with and without a vertical support motion of the base nodes
(ComputeSupportMotionVector, the base DOFs read through svlg_get_displacements), each against
the oracle to 1e-10 relative L2; it exits 0 when every check passes."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "capi" / "svlg_abi_check"


def test_compiled_caller_matches_oracle():
    if not BIN.exists():
        raise AssertionError(f"{BIN} is not built; run __graft_entry__.build() first")
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
