"""GPU checks of the benchmark workloads at (or near) their real sizes.

Against the oracle (relative L2 <= 1e-10), with references that take the oracle minutes written
once to tests/golden/config_refs.npz by tests/golden/make_config_refs.py:
  * the 2D layer (lin2DQuad4 200 x 100, 40,602 free DOFs) over all 2000 CentralDifference steps,
    recorded DOFs and the final field;
  * the layered, lognormal-Vs half-space at 24 x 24 x 12 elements over 4000 steps (drift over
    the full run length of the 3D workload), and its whole field after step 1000;
  * each of the 64 batched realizations (3.4 M DOFs in one union) over 50 steps;
  * the full-size 3D half-space (100 x 100 x 50, 1,560,753 free DOFs, 4,436 node tiles with the
    compile-time slot stride) over its first 3 steps against a live oracle run, with the source
    moved early so that the three steps carry force;
  * SVL's F02, F06 and J05 validation cases on the device Newmark path against the OpenSees
    histories SVL ships (6 significant digits: 1e-5 relative RMS, as the oracle's own test).
Size-independent properties at full size: bitwise run-to-run determinism and split calls = one
call (half-space, 200 steps); superposition of two loads (relative 1e-12); on the 64^3 implicit
block and the 80 x 80 x 40 J2 block, the CG residual recomputed by an independent svlg_keff_apply
and bitwise determinism of the Newton steps.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from svl_amd import batch as B
from svl_amd import workloads as W
from svl_amd.integrators import CentralDifference, DeviceModel, ExtendedNewmarkBeta, LinearSystem, NewmarkBeta

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
REL_TOL = 1e-10


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def refs():
    with np.load(GOLD / "config_refs.npz") as z:
        return {k: z[k] for k in z.files}


def signature(m):
    return np.array([m.coords.sum(), np.sum(m.E), np.sum(m.dash_eta), np.sum(m.load_series), m.nfree, m.nelem])


def cd_run(m, dt, nsteps, rec, **opts):
    dm = DeviceModel(m, **opts)
    cd = CentralDifference(dm, dt)
    cd.Initialize()
    hist = cd.advance(1, nsteps, rec)
    U = cd.GetDisplacements()
    dm.close()
    return hist, U


def test_config1_layer_all_2000_steps():
    g = refs()
    m, dt, nt = W.layer2d(200, 100, nt=2001)
    assert np.allclose(signature(m), g["c1_sig"], rtol=1e-14, atol=0), "workload generator changed"
    hist, U = cd_run(m, dt, nt - 1, g["c1_rec"])
    e1, e2 = rel(hist, g["c1_U"]), rel(U, g["c1_final"])
    print(f"2D layer 200x100, 2000 steps: history {e1:.2e}, final field {e2:.2e}")
    assert e1 < REL_TOL and e2 < REL_TOL


def test_config2_shaped_halfspace_4000_steps():
    g = refs()
    m, dt, nt = W.halfspace3d(24, 24, 12, nt=4001)
    assert np.allclose(signature(m), g["c2s_sig"], rtol=1e-14, atol=0), "workload generator changed"
    hist, U = cd_run(m, dt, nt - 1, g["c2s_rec"])
    _, U1000 = cd_run(m, dt, 1000, [])
    e1, e2 = rel(hist[3::4], g["c2s_U"]), rel(U1000, g["c2s_mid"])
    print(f"half-space 24x24x12: 4000-step history {e1:.2e}, field after step 1000 {e2:.2e}")
    assert e1 < REL_TOL and e2 < REL_TOL


def test_config4_all_64_realizations():
    g = refs()
    models, dts, nt = W.batch2d(nt=51)
    for i, m in enumerate(models):
        assert np.allclose(signature(m), g["c4_sig"][i], rtol=1e-14, atol=0), "workload generator changed"
    u, nptr, fptr = B.union(models)
    dm = DeviceModel(u)
    cd = B.BatchedCentralDifference(dm, dts, nptr, fptr)
    cd.Initialize()
    rec = np.concatenate([fptr[i] + g["c4_rec"][i] for i in range(len(models))]).astype(np.int32)
    hist = cd.advance(1, nt - 1, rec_dofs=rec)
    dm.close()
    worst = 0.0
    for i in range(len(models)):
        e = rel(hist[:, 4 * i:4 * i + 4], g["c4_U"][i])
        worst = max(worst, e)
        assert e < REL_TOL, (i, e)
    print(f"64 realizations ({u.nfree} DOFs), 50 steps: worst part rel L2 {worst:.2e}")


def test_config2_full_size_first_steps():
    """The full 100 x 100 x 50 half-space (every tile of the production plan) against a live
    oracle run of three steps; the Ricker source is centred at 2 dt so the steps carry force."""
    m, dt, nt = W.halfspace3d(100, 100, 50, nt=4)
    m, dt, nt = W.halfspace3d(100, 100, 50, nt=4, t0=2 * dt)
    src = W.dofs_of(m, int(m.load_node[0]))
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, src)
    hist, U = cd_run(m, dt, nt - 1, src)
    dm = DeviceModel(m)
    info = dm.info()
    dm.close()
    e1, e2 = rel(hist, ref["U"]), rel(U, ref["final_U"])
    print(f"half-space 100x100x50 ({info['ntiles']} tiles, path {info['elem_path']}): 3 steps, "
          f"source history {e1:.2e}, final field {e2:.2e} ({np.count_nonzero(ref['final_U'])} nonzero DOFs)")
    assert np.count_nonzero(ref["final_U"]) > 100
    assert e1 < REL_TOL and e2 < REL_TOL


def test_config2_full_size_determinism_split_and_superposition():
    m, dt, nt = W.halfspace3d(100, 100, 50, nt=401)
    src = int(m.load_node[0])
    # A second load: horizontal, at a node a quarter of the way along x.
    other = m.nearest_node((-50.0, 0.0, 0.0))
    s2 = np.sin(np.arange(nt) * 0.05) * 2e5
    rec = W.dofs_of(m, src) + W.dofs_of(m, other) + list(range(11, m.nfree, 100003))
    dm = DeviceModel(m)
    cd = CentralDifference(dm, dt)
    runs = []
    for split in (False, True, False):
        cd.Initialize()
        if split:
            h = np.concatenate([cd.advance(1, 137, rec), cd.advance(138, 63, rec)])
        else:
            h = cd.advance(1, 200, rec)
        runs.append((h, cd.GetDisplacements()))
    dm.close()
    assert np.array_equal(runs[0][0], runs[2][0]) and np.array_equal(runs[0][1], runs[2][1])
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    # Superposition: U(load A + load B) = U(A) + U(B).
    mb = W.halfspace3d(100, 100, 50, nt=401)[0]
    mb.load_node, mb.load_dir, mb.load_factor = mb.load_node[:0], np.zeros((0, 3)), np.zeros(0)
    mb.load_series = np.zeros((0, nt))
    mb.add_point_load(other, [1.0, 0.0, 0.0], s2)
    mab = W.halfspace3d(100, 100, 50, nt=401)[0]
    mab.add_point_load(other, [1.0, 0.0, 0.0], s2)
    _, Ub = cd_run(mb, dt, 200, rec)
    _, Uab = cd_run(mab, dt, 200, rec)
    e = rel(runs[0][1] + Ub, Uab)
    print(f"half-space 100x100x50, 200 steps: determinism and split calls bitwise; superposition {e:.2e}")
    assert np.abs(Ub).max() > 0 and e < 1e-12


def test_config3_full_size_cg_residual_recomputed():
    """The 64^3 implicit block: after 5 steps, the effective force of step 6 is solved with
    svlg_solve at 1e-8 and the residual is recomputed with svlg_keff_apply (an operator launch
    independent of the solver's own partial sums)."""
    m, dt, nt = W.block3d_implicit(64, 64, 64, nt=8)
    dm = DeviceModel(m, tile_nodes=330)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-8, maxit=2000)
    nm.Initialize()
    nm.advance(1, 5)
    b = nm.ComputeEffectiveForce(6)
    ls = LinearSystem(nm)
    ls.SetVector(b)
    assert not ls.SolveSystem()
    x = ls.GetSolution()
    r = b - ls.Apply(x)
    res = float(np.linalg.norm(r) / np.linalg.norm(b))
    print(f"64^3 block ({m.nfree} DOFs): CG {ls.iters} iterations, reported {ls.relres:.2e}, recomputed {res:.2e}")
    assert res <= 1.05e-8 and abs(res - ls.relres) <= 0.05 * ls.relres
    dm.close()


def test_config5_full_size_newton_determinism():
    m, dt, nt = W.soil_structure_j2(80, 80, 40, nt=161)  # 154 steps: into the strong part of the load
    out = []
    for _ in range(2):
        dm = DeviceModel(m)
        nm = NewmarkBeta(dm, dt, cnvgtol=1e-6, nmax=20, cnvgtest=5, rtol=1e-8, maxit=3000)
        nm.Initialize()
        nm.advance(1, 154)
        out.append((nm.GetDisplacements(), nm.iters.copy()))
        dm.close()
    print(f"J2 soil-structure 80x80x40 ({m.nfree} DOFs): Newton iterations of the last steps {out[0][1][-6:]}, bitwise "
          f"{np.array_equal(out[0][0], out[1][0])}")
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.all(np.isfinite(out[0][0])) and np.abs(out[0][0]).max() > 0


def _fixture(name):
    with np.load(GOLD / f"{name}.npz") as z:
        return {k: z[k] for k in z.files}


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


@pytest.mark.parametrize("case", ["F02", "F06", "J05"])
def test_svl_validation_cases_on_device(case):
    """SVL's elastic validation cases on the device Newmark path (the models as rebuilt by the
    oracle's fixture test) against the OpenSees histories SVL ships."""
    import test_oracle_fixtures as T
    g = _fixture(case)
    if case == "F02":
        from svl_amd.mesh import ELEM_QUAD4, Model, tags_for
        coords = np.array([[0.0, 0.0], [2.0, 0.0], [0.0, 2.0], [2.0, 2.0]])
        restrained = np.array([[True, True], [False, True], [False, False], [False, False]])
        m = Model(ndim=2, coords=coords, node_tags=tags_for(4), elem_type=ELEM_QUAD4, conn=np.array([[0, 1, 3, 2]]),
                  E=2.08e8, nu=0.30, rho=2000.0, thickness=1.0, restrained=restrained)
        m.add_point_load(3, [0.01, -0.02], g["load"])
        rec, dt, nt, ref_cols = [int(d) for d in m.free_dof()[3]], 0.025, 201, [1, 2]
    elif case == "F06":
        m = T.soil_column_2d(g)
        rec, dt, nt, ref_cols = [int(m.free_dof()[200][0])], 0.01, 1000, [3]
    else:
        from svl_amd.mesh import ELEM_HEXA8, Model, make_domain_volume, tags_for
        coords, conn, bnd = make_domain_volume(1, 1, 100, (0, 0, 0), (1.0, 0, 0), (0, 1.0, 0), (0, 0, 100.0))
        restrained = np.zeros(coords.shape, bool)
        restrained[:, 2] = True
        m = Model(ndim=3, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_HEXA8, conn=conn,
                  E=1.3e7, nu=0.30, rho=2000.0, restrained=restrained, ray_am=0.1244195110, ray_ak=0.0007878958,
                  dash_node=np.repeat(np.arange(4), 2), dash_dir=np.tile([0, 1], 4), dash_eta=np.full(8, 2.5e4))
        for n in range(4):
            m.add_point_load(n, [25000.0, 0.0, 0.0], g["load"])
        rec, dt, nt, ref_cols = [int(m.free_dof()[400][0])], 0.01, 1000, [3]
    ref = oracle.run(m, oracle.NEWMARK, dt, nt, rec)
    dm = DeviceModel(m)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=20000)
    nm.Initialize()
    hist = nm.advance(1, nt - 1, rec)
    dm.close()
    os_u = g["displacement"][:, ref_cols]
    e_or, e_os = rel(hist, ref["U"]), _rel_rms(hist, os_u)
    print(f"{case} on the device: vs oracle {e_or:.2e}, vs OpenSees {e_os:.2e} (relative RMS)")
    assert e_or < REL_TOL and e_os < 1e-5
