"""GPU parity of NewmarkBeta + NewtonRaphson with J2 elements (svlg_nl_*).

The oracle restates SVL's Newton loop with a direct (skyline LDL^T) solve of the tangent system
per iteration; the device solves it by Jacobi-CG to 1e-13, so that residual is the only
difference. Bars: displacement histories to relative L2 <= 1e-10 with the Newton iteration
count of every step equal to the oracle's, and the convergence-test value of every iteration
to 1e-6 relative plus a round-off floor (1e-9 plus 1e-9 of the step's first value).

One branch of SVL's algorithm is decided by round-off: the state update at the converged
iterate of a yielding step (Algorithm::UpdateStatesIncrements after the last increment)
evaluates the yield function of a point that the previous update left on the surface, so
F = +-round-off and the point keeps the elastoplastic tangent or falls back to the elastic one
depending on the last bits. The next step then starts from different tangents and takes
Newton iterations 2+1 or 1+2 across two steps; both are SVL's algorithm. Cases where this
happens (the F07 soil column) assert the exact bars up to the first such step and, after it,
a history difference below 10x the Newton tolerance (the answers differ by converged Newton
corrections, which the iterate accumulation carries forward), with differing iteration counts
on at most 2% of the steps.
An elastic model through Newton reproduces the Linear ExtendedNewmarkBeta history.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from svl_amd import workloads as W
from svl_amd.integrators import DeviceModel, ExtendedNewmarkBeta, LearnedPreconditioner, NewmarkBeta
from svl_amd.mesh import ELEM_QUAD4, Model, tags_for

from _meshes import block_model, rotation

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_device(m, dt, nt, rec, cnvgtol, nmax, cnvgtest=5):
    dm = DeviceModel(m)
    nm = NewmarkBeta(dm, dt, cnvgtol=cnvgtol, nmax=nmax, cnvgtest=cnvgtest, rtol=1e-13, maxit=20000)
    nm.Initialize()
    hist = nm.advance(1, nt - 1, rec)
    return dm, nm, hist


def check(m, dt, nt, rec, cnvgtol, nmax, cnvgtest=5, branch_free=True):
    print(f"nonlinear case: ndim={m.ndim} nelem={m.nelem} nt={nt} test={cnvgtest}", flush=True)
    ref = oracle.run(m, oracle.NEWMARK, dt, nt, rec, algorithm=oracle.NEWTON, nmax=nmax, cnvgtest=cnvgtest,
                     cnvgtol=cnvgtol)
    dm, nm, hist = run_device(m, dt, nt, rec, cnvgtol, nmax, cnvgtest)
    err = rel_l2(hist, ref["U"])
    diff = np.nonzero(nm.iters != ref["iters"])[0]
    first = int(diff[0]) if diff.size else nt - 1
    err_before = rel_l2(hist[:first], ref["U"][:first]) if first > 0 else 0.0
    print(f"  rel L2 {err:.2e}; Newton iterations differ on {diff.size} of {nt - 1} steps (first at {first}); "
          f"rel L2 before it {err_before:.2e}; max iterations {int(ref['iters'].max())}", flush=True)
    for k in diff[:3]:
        print(f"  step {k} residuals device {np.array2string(nm.residuals[k][:5], precision=6)} "
              f"oracle {np.array2string(ref['residuals'][k][:5], precision=6)}", flush=True)
    assert ref["iters"].max() > 2  # the material yields and Newton iterates
    # Convergence-test values of the steps before the first differing one.
    # Round-off floor: 1e-9 plus 1e-9 of the step's first value (the values of one step share
    # units; below 1e-9 the relative tests 5 and 7 read the solves' round-off).
    a, b = nm.residuals[:first], ref["residuals"][:first]
    assert np.array_equal(np.isnan(a), np.isnan(b))
    floor = 1e-9 * np.abs(b[:, :1]) + 1e-9
    ok = np.isnan(b) | (np.abs(a - b) <= 1e-6 * np.abs(b) + floor)
    bad = np.argwhere(~ok)[:5]
    assert np.all(ok), [(int(k), int(i), a[k, i], b[k, i]) for k, i in bad]
    assert err_before < 1e-10
    if branch_free:
        assert diff.size == 0
        assert err < 1e-10
    else:
        assert diff.size <= 0.02 * (nt - 1)
        assert err < 10 * cnvgtol
    nm.oracle_ref = ref
    return dm, nm


def yield_check(nm, m):
    """Gauss-point stresses are finite and some J2 point has reached the yield surface:
    ||dev sigma|| >= sqrt(2/3) sigma_y (3D), |tau_xy| >= sigma_y / sqrt(3) in the simple
    shear of the soil column (2D)."""
    reached = False
    for e in np.nonzero(m.mat_type == 1)[0]:
        s = nm.GetStress(int(e))
        assert np.all(np.isfinite(s))
        sy = m.j2[e][2]
        if m.ndim == 2:
            reached |= bool(np.any(np.abs(s[:, 2]) >= (1 - 1e-9) * sy / np.sqrt(3.0)))
        else:
            d = s.copy()
            d[:, :3] -= s[:, :3].mean(1)[:, None]
            dev = np.sqrt((d[:, :3] ** 2).sum(1) + 2 * (d[:, 3:] ** 2).sum(1))
            reached |= bool(np.any(dev >= (1 - 1e-9) * np.sqrt(2 / 3) * sy))
    assert reached


def test_single_quad4_j2_F03_load():
    """One plane-strain J2 Quad4 under SVL's F03 load (isotropic hardening)."""
    with np.load(GOLD / "F03.npz") as z:
        load = z["load"]
    coords = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0]])
    restrained = np.array([[True, True], [False, True], [False, True], [False, True]])
    m = Model(ndim=2, coords=coords, node_tags=tags_for(4), elem_type=ELEM_QUAD4, conn=np.array([[0, 1, 2, 3]]),
              E=1.0, nu=0.3, rho=0.0, thickness=1.0, restrained=restrained)
    m.set_j2(1.333333e8, 8.0e7, 4.0e7, 8.0e7, 1.0)
    m.add_point_load(2, [8.0e7, 0.0], load)
    nt = 600
    rec = [int(m.free_dof()[2][0])]
    dm, nm = check(m, 0.001, nt, rec, 1e-6, 100)
    dm.close()


def soil_column(n, nt, beta, k0=0):
    """F07's column with n node pairs, driven by the F07 load from sample k0 on."""
    with np.load(GOLD / "F07.npz") as z:
        load = z["load"][k0:k0 + nt]
    coords = np.zeros((2 * n, 2))
    coords[0::2, 0], coords[1::2, 0] = 0.0, 1.0
    coords[0::2, 1] = coords[1::2, 1] = np.arange(n, dtype=float)
    k = np.arange(n - 1)
    conn = np.stack([2 * k, 2 * k + 1, 2 * k + 3, 2 * k + 2], axis=1).astype(np.int32)
    restrained = np.zeros(coords.shape, bool)
    restrained[:, 1] = True
    m = Model(ndim=2, coords=coords, node_tags=tags_for(2 * n), elem_type=ELEM_QUAD4, conn=conn, E=1.0, nu=0.3,
              rho=2000.0, thickness=1.0, restrained=restrained, ray_am=0.0, ray_ak=0.0005,
              dash_node=np.array([0, 1]), dash_dir=np.array([0, 0]), dash_eta=np.array([1.0e6, 1.0e6]))
    m.add_point_load(0, [1.0e6, 0.0], load)
    m.add_point_load(1, [1.0e6, 0.0], load)
    m.set_j2(2.9e7, 2.0e7, 1.0e4, 1.0e7, beta)
    return m


@pytest.mark.parametrize("beta", [1.0, 0.0])
def test_soil_column_F07(beta):
    """SVL's F07 J2 soil column (shortened to 40 elements and the strong part of the load,
    samples 600-1000), isotropic and kinematic hardening."""
    nt = 401
    m = soil_column(41, nt, beta, k0=600)
    rec = [int(m.free_dof()[0][0]), int(m.free_dof()[40][0]), int(m.free_dof()[81][0])]
    dm, nm = check(m, 0.001, nt, rec, 1e-6, 50, branch_free=False)
    yield_check(nm, m)
    dm.close()


def j2_block(test=5, transform=None):
    m, dt, nt, src = block_model(4, 4, 3, transform=transform)
    m.ray_am = np.full(m.nelem, 0.5)
    m.ray_ak = np.full(m.nelem, 1e-3)
    sel = np.nonzero(np.arange(m.nelem) % 2 == 0)[0]
    E, nu = np.asarray(m.E)[sel], 0.3
    K, G = E / 3 / (1 - 2 * nu), E / 2 / (1 + nu)
    m.set_j2(K, G, 2e5, 0.1 * G, 0.5, sel)
    return m, dt, src


@pytest.mark.parametrize("cnvgtest,tol", [(5, 1e-8), (1, 1e-3), (7, 1e-9)])
def test_hexa8_block_mixed_j2_elastic(cnvgtest, tol):
    m, dt, src = j2_block()
    nt = 100
    rec = W.dofs_of(m, src) + list(range(0, m.nfree, 17))
    dm, nm = check(m, dt, nt, rec, tol, 30, cnvgtest)
    yield_check(nm, m)
    dm.close()


def test_hexa8_block_mixed_j2_rotated():
    """The same mixed block rotated and sheared (parallelepipeds off the axes): the J2 points
    take the Gauss loop on the element coordinates instead of the axis-aligned corner sums."""
    m, dt, src = j2_block(transform=rotation())
    nt = 60
    rec = W.dofs_of(m, src) + list(range(0, m.nfree, 17))
    dm, nm = check(m, dt, nt, rec, 1e-8, 30, 5)
    assert dm.info()["elem_path"] < 3
    dm.close()


def test_elastic_newton_equals_linear():
    m, dt, nt, src = block_model(5, 4, 3)
    m.ray_am = np.full(m.nelem, 0.5)
    m.ray_ak = np.full(m.nelem, 1e-3)
    nt = 60
    rec = W.dofs_of(m, src)
    dm, nm, hist = run_device(m, dt, nt, rec, 1e-9, 10)
    dm.close()
    dm2 = DeviceModel(m)
    lin = ExtendedNewmarkBeta(dm2, dt, rtol=1e-13, maxit=5000)
    lin.Initialize()
    h2 = lin.advance(1, nt - 1, rec)
    dm2.close()
    # Newton's second iterate finds no correction beyond the CG tolerance (one iterate where
    # the effective force vanishes).
    assert np.all(nm.iters <= 2) and np.sum(nm.iters == 2) > nt // 2
    assert rel_l2(hist, h2) < 1e-10


def test_soil_structure_jacobi_and_learned():
    """Small soil-structure block (J2 soil layers around a stiff elastic block, strong
    horizontal pulse): Jacobi-CG Newton against the oracle, then the same run with the
    learned (Chebyshev-initialised, BF16 MFMA) preconditioner in flexible CG, which approximates
    the initial effective stiffness while CG solves the tangent systems."""
    m, dt, nt = W.soil_structure_j2(8, 8, 4, nt=41, block=(2, 2, 2), Ap=2e7, t0=0.08, f0=8.0)
    rec = W.dofs_of(m, m.nearest_node((0.0, 0.0, 0.0))) + list(range(0, m.nfree, 23))
    dm, nm = check(m, dt, nt, rec, 1e-6, 20, 5, branch_free=False)
    yield_check(nm, m)
    it_j, cg_j = nm.iters.copy(), nm.cg_iters.copy()
    dm.close()
    ref = nm.oracle_ref
    dm = DeviceModel(m)
    nl = NewmarkBeta(dm, dt, cnvgtol=1e-6, nmax=20, cnvgtest=5, rtol=1e-13, maxit=5000)
    nl.Initialize()
    pc = LearnedPreconditioner(nl, channels=1, layers=6, mfma=-1)
    assert pc.info()["mfma"] == 1
    pc.use(True)
    hist = nl.advance(1, nt - 1, rec)
    print(f"  CG iterations per Newton solve: Jacobi {cg_j.sum() / it_j.sum():.1f}, learned "
          f"{nl.cg_iters.sum() / nl.iters.sum():.1f}", flush=True)
    assert rel_l2(hist, ref["U"]) < 1e-6
    assert np.mean(nl.iters != ref["iters"]) <= 0.05
    assert nl.cg_iters.sum() / nl.iters.sum() < 0.7 * cg_j.sum() / it_j.sum()
    dm.close()
