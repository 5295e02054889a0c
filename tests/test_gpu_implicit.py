"""GPU parity of the implicit ExtendedNewmarkBeta path (matrix-free Keff, Jacobi-PCG).

Bars:
  * Keff x against the oracle's assembled T^T (K + 4/dt^2 M + 2/dt C) T: relative L2
    <= 1e-13 (same arithmetic up to summation order);
  * displacement histories against the oracle's direct (skyline LDL^T) solves, as SVL's
    EigenSolver does them: relative L2 <= 1e-10 with the CG tolerance at 1e-13, so the
    CG residual is the only difference; final V and A to the same bar;
  * svlg_solve: the oracle's operator applied to the returned x reproduces b to the
    requested tolerance (times a small factor for the round-off of the two operators).
"""
import numpy as np
import pytest

import oracle
from svl_amd import workloads as W
from svl_amd._capi import SvlgError
from svl_amd.integrators import DeviceModel, ExtendedNewmarkBeta, LinearSystem

from _meshes import block_model, graded, rotation

pytestmark = pytest.mark.gpu

REL_TOL = 1e-10


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def with_rayleigh(m, am=0.5712, ak=1.447e-3):
    m.ray_am = np.full(m.nelem, am)
    m.ray_ak = np.full(m.nelem, ak)
    return m


def hexa_case(kind):
    if kind == "axis":
        m, dt, nt, src = block_model(7, 6, 5)
    elif kind == "rotated":
        m, dt, nt, src = block_model(7, 6, 5, transform=rotation())
    elif kind == "graded":
        m, dt, nt, src = block_model(7, 6, 5)
        m = graded(m)
    else:
        m, dt, nt, src = block_model(7, 6, 5, distort=0.18, seed=4)
    return with_rayleigh(m), dt, nt


def newmark(dm, dt, rtol=1e-13):
    nm = ExtendedNewmarkBeta(dm, dt, rtol=rtol, maxit=5000)
    nm.Initialize()
    return nm


@pytest.mark.parametrize("kind,geo,path", [("axis", 0, 3), ("graded", 0, 3), ("rotated", 0, 2), ("axis", -1, 1),
                                           ("distorted", 0, 0)])
@pytest.mark.parametrize("tile_nodes", [24, 300])
def test_keff_apply_hexa(kind, geo, path, tile_nodes):
    m, dt, _ = hexa_case(kind)
    dt_nm = 1e-3
    dm = DeviceModel(m, tile_nodes=tile_nodes, geo_table=geo)
    assert dm.info()["elem_path"] == path and dm.info()["variant"] == 3
    newmark(dm, dt_nm)
    ls = LinearSystem(ExtendedNewmarkBeta(dm, dt_nm))
    rng = np.random.default_rng(7)
    for _ in range(2):
        x = rng.standard_normal(m.nfree)
        y = ls.Apply(x)
        assert rel_l2(y, oracle.keff_apply(m, dt_nm, x)) < 1e-13
    dm.close()


@pytest.mark.parametrize("geo", [0, -1])
def test_keff_apply_quad(geo):
    m, dt, nt = W.layer2d(24, 12, nt=101)
    m = with_rayleigh(m, 0.3, 2e-3)
    dm = DeviceModel(m, tile_nodes=60, geo_table=geo)
    newmark(dm, 4 * dt)
    x = np.random.default_rng(3).standard_normal(m.nfree)
    y = LinearSystem(ExtendedNewmarkBeta(dm, 4 * dt)).Apply(x)
    assert rel_l2(y, oracle.keff_apply(m, 4 * dt, x)) < 1e-13
    dm.close()


@pytest.mark.parametrize("kind", ["axis", "rotated", "distorted"])
def test_newmark_history_hexa(kind):
    m, dt, nt = hexa_case(kind)
    dt, nt = 1e-3, 120
    rec = list(range(0, m.nfree, 9))
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec, want_va=True)
    dm = DeviceModel(m, tile_nodes=80)
    nm = newmark(dm, dt)
    hist = nm.advance(1, nt - 1, rec)
    assert np.all(nm.relres <= 1e-13) and np.all(nm.iters >= 0)
    assert rel_l2(hist, ref["U"]) < REL_TOL
    U, V, A = nm._state()
    assert rel_l2(U, ref["final_U"]) < REL_TOL
    assert rel_l2(V[rec], ref["V"][-1]) < REL_TOL
    assert rel_l2(A[rec], ref["A"][-1]) < REL_TOL
    dm.close()


def test_newmark_history_quad_layer():
    m, dt, nt = W.layer2d(24, 12, nt=301)
    m = with_rayleigh(m, 0.3, 2e-3)
    rec = W.dofs_of(m, int(m.load_node[0])) + list(range(3, m.nfree, 17))
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec)
    dm = DeviceModel(m, tile_nodes=100)
    nm = newmark(dm, dt)
    hist = nm.advance(1, nt - 1, rec)
    assert rel_l2(hist, ref["U"]) < REL_TOL
    dm.close()


def test_newmark_block_workload_small():
    """The implicit block workload (random Vs field, Rayleigh 5 % at 1 and 10 Hz, fixed
    base) at 8^3 elements, stepped one call at a time as SVL's DynamicAnalysis does."""
    m, dt, nt = W.block3d_implicit(8, 8, 8, nt=61)
    rec = W.dofs_of(m, int(m.load_node[0])) + list(range(0, m.nfree, 23))
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec)
    dm = DeviceModel(m)
    nm = newmark(dm, dt)
    rows = []
    for k in range(1, nt):
        assert not nm.ComputeNewStep(k)
        rows.append(nm.GetDisplacements()[rec])
    assert rel_l2(np.array(rows), ref["U"]) < REL_TOL
    dm.close()


def test_solve_residual_and_zero_rhs():
    m, dt, nt = hexa_case("rotated")
    dm = DeviceModel(m, tile_nodes=64)
    nm = newmark(dm, 1e-3, rtol=1e-11)
    ls = LinearSystem(nm)
    b = np.random.default_rng(11).standard_normal(m.nfree) * 1e4
    ls.SetVector(b)
    assert not ls.SolveSystem()
    x = ls.GetSolution()
    assert ls.relres <= 1e-11 and ls.iters > 0
    assert rel_l2(oracle.keff_apply(m, 1e-3, x), b) < 1e-10
    ls.SetVector(np.zeros(m.nfree))
    assert not ls.SolveSystem()
    assert ls.iters == 0 and not np.any(ls.GetSolution())
    dm.close()


def test_newmark_deterministic():
    m, dt, nt = W.block3d_implicit(10, 9, 7, nt=41)
    out = []
    for _ in range(2):
        dm = DeviceModel(m, tile_nodes=120)
        nm = newmark(dm, dt, rtol=1e-10)
        nm.advance(1, nt - 1)
        out.append((nm.GetDisplacements(), nm.iters.copy()))
        dm.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_newmark_rejects_unsupported_models():
    m, dt, nt = hexa_case("axis")
    dm = DeviceModel(m, variant=1)
    with pytest.raises(SvlgError):
        newmark(dm, 1e-3)
    dm.close()
