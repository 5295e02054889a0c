"""GPU: svlg_keff_apply_noalloc (y = Keff x with the work vectors in the library's fixed device
array, no allocation in the call) against svlg_keff_apply, through the C ABI.

Bars: the two entry points run the same kernels, so their outputs are bit-identical; both
match the oracle's assembled effective stiffness to 1e-13 relative. Two calls in a row on one
model, then a call on a smaller model (the array keeps the previous call's contents, which the
call must not read). A model with more node DOFs than the array holds is refused with an error
that says so, before any device work.
"""
import numpy as np
import pytest

import oracle
from svl_amd import _capi as capi
from svl_amd import workloads as W
from svl_amd.integrators import DeviceModel, ExtendedNewmarkBeta, LinearSystem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def noalloc(dm, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(dm.nfree)
    capi.check(capi.lib().svlg_keff_apply_noalloc(dm.h, capi.fptr(x), capi.fptr(y)), "svlg_keff_apply_noalloc")
    return y


def setup(nx, ny, nz, consistent=False):
    m, dt, nt = W.block3d_implicit(nx, ny, nz, nt=3)
    m.massform_lumped = not consistent
    dm = DeviceModel(m)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-10)
    nm.Initialize()
    return m, dt, dm, LinearSystem(nm)


def test_noalloc_matches_keff_apply_twice_then_smaller():
    rng = np.random.default_rng(31)
    big = setup(12, 11, 9)
    small = setup(5, 4, 6, consistent=True)
    try:
        for (m, dt, dm, ls), reps in ((big, 2), (small, 1)):
            for r in range(reps):
                x = rng.standard_normal(m.nfree)
                y_ref = ls.Apply(x)
                y = noalloc(dm, x)
                print(f"noalloc nfree={m.nfree} call {r}: identical={np.array_equal(y, y_ref)}, "
                      f"vs oracle {rel(y, oracle.keff_apply(m, dt, x)):.2e}", flush=True)
                assert np.array_equal(y, y_ref)
                assert rel(y, oracle.keff_apply(m, dt, x)) < 1e-13
    finally:
        big[2].close()
        small[2].close()


def test_noalloc_refuses_models_beyond_the_array():
    # 70^3 Hexa8 block: 71^3 nodes x 3 = 1,073,733 node DOFs > 2^20.
    m, dt, nt = W.block3d_implicit(70, 70, 70, nt=3)
    assert m.nnodes * m.ndim > 1 << 20
    dm = DeviceModel(m)
    try:
        x = np.zeros(m.nfree)
        with pytest.raises(capi.SvlgError, match="scratch array"):
            noalloc(dm, x)
    finally:
        dm.close()
