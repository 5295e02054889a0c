"""GPU parity of the implicit paths with SVL's consistent element mass (massform CONSISTENT,
SVL's default: Driver.hpp:1998; lin3DHexa8.cpp / lin2DQuad4.cpp:243-285 before the lumping
branch).

The device applies M_e element by element inside the effective-stiffness operator (modal form
on parallelepipeds, Gauss points otherwise), in the Newmark right-hand side (M_e (4/dt V + A +
am V)), in the Jacobi diagonal, in the Newton residual, in the network's operator and in the
reactions. Bars as for the lumped paths:
  * Keff x against the oracle's assembled T^T (K + 4/dt^2 M + 2/dt C) T: relative L2 <= 1e-13;
  * histories against the oracle's direct solves: relative L2 <= 1e-10 with CG at 1e-13;
  * SVL's J02 case (one Hexa8, consistent mass, Newmark) against the oracle to 1e-10 and the
    OpenSees history SVL ships to 1e-5 (6 significant digits);
  * Newton with J2 elements: the exact bars of test_gpu_nonlinear (iteration counts, values);
  * reactions: relative L2 <= 1e-9.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from svl_amd import workloads as W
from svl_amd._capi import SvlgError
from svl_amd.integrators import CentralDifference, DeviceModel, ExtendedNewmarkBeta, LearnedPreconditioner, LinearSystem
from svl_amd.mesh import ELEM_HEXA8, Model, tags_for

from _meshes import block_model, graded, rotation

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def consistent(m, am=0.5712, ak=1.447e-3):
    m.massform_lumped = False
    m.ray_am = np.full(m.nelem, am)
    m.ray_ak = np.full(m.nelem, ak)
    return m


def hexa_case(kind):
    if kind == "axis":
        m = block_model(6, 5, 4)[0]
    elif kind == "rotated":
        m = block_model(6, 5, 4, transform=rotation())[0]
    elif kind == "graded":
        m = graded(block_model(6, 5, 4)[0])
    else:
        m = block_model(6, 5, 4, distort=0.18, seed=4)[0]
    return consistent(m)


def newmark(dm, dt, rtol=1e-13):
    nm = ExtendedNewmarkBeta(dm, dt, rtol=rtol, maxit=5000)
    nm.Initialize()
    return nm


@pytest.mark.parametrize("kind,geo,path", [("axis", 0, 3), ("graded", 0, 3), ("rotated", 0, 2), ("axis", -1, 1),
                                           ("distorted", 0, 0)])
def test_keff_apply_hexa_consistent(kind, geo, path):
    m = hexa_case(kind)
    dt = 1e-3
    dm = DeviceModel(m, tile_nodes=40, geo_table=geo)
    assert dm.info()["elem_path"] == path
    ls = LinearSystem(newmark(dm, dt))
    rng = np.random.default_rng(5)
    for _ in range(2):
        x = rng.standard_normal(m.nfree)
        err = rel(ls.Apply(x), oracle.keff_apply(m, dt, x))
        print(f"Keff x {kind} path {path}: rel {err:.2e}")
        assert err < 1e-13
    dm.close()


@pytest.mark.parametrize("geo,distort", [(0, 0.0), (-1, 0.0), (0, 0.2)])
def test_keff_apply_quad_consistent(geo, distort):
    m, dt, nt = W.layer2d(20, 10, nt=51)
    if distort:
        rng = np.random.default_rng(2)
        inner = np.nonzero(~m.restrained.any(1) & (m.coords[:, 1] < m.coords[:, 1].max()))[0]
        c = m.coords.copy()
        c[inner] += rng.uniform(-distort, distort, (inner.size, 2))
        m.coords = np.ascontiguousarray(c)
    m = consistent(m, 0.3, 2e-3)
    dm = DeviceModel(m, tile_nodes=60, geo_table=geo)
    ls = LinearSystem(newmark(dm, 4 * dt))
    x = np.random.default_rng(3).standard_normal(m.nfree)
    err = rel(ls.Apply(x), oracle.keff_apply(m, 4 * dt, x))
    print(f"Keff x quad geo {geo} distort {distort} path {dm.info()['elem_path']}: rel {err:.2e}")
    assert err < 1e-13
    dm.close()


@pytest.mark.parametrize("kind", ["axis", "rotated", "distorted"])
def test_newmark_history_hexa_consistent(kind):
    m = hexa_case(kind)
    dt, nt = 1e-3, 100
    rec = list(range(0, m.nfree, 7))
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec, want_va=True)
    dm = DeviceModel(m, tile_nodes=80)
    nm = newmark(dm, dt)
    hist = nm.advance(1, nt - 1, rec)
    U, V, A = nm._state()
    errs = (rel(hist, ref["U"]), rel(U, ref["final_U"]), rel(V[rec], ref["V"][-1]), rel(A[rec], ref["A"][-1]))
    print(f"consistent-mass Newmark {kind}: U hist {errs[0]:.2e} final U {errs[1]:.2e} V {errs[2]:.2e} "
          f"A {errs[3]:.2e}, CG iterations {nm.iters.mean():.1f}")
    assert max(errs) < 1e-10
    dm.close()


def test_newmark_history_quad_consistent():
    m, dt, nt = W.layer2d(24, 12, nt=201)
    m = consistent(m, 0.3, 2e-3)
    rec = W.dofs_of(m, int(m.load_node[0])) + list(range(3, m.nfree, 17))
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec)
    dm = DeviceModel(m, tile_nodes=100)
    hist = newmark(dm, dt).advance(1, nt - 1, rec)
    err = rel(hist, ref["U"])
    print(f"consistent-mass Newmark Quad4 layer: rel L2 {err:.2e}")
    assert err < 1e-10
    dm.close()


def j02_model(g):
    coords = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2]],
                      dtype=float)
    restrained = np.zeros((8, 3), bool)
    restrained[0] = True
    restrained[1:4, 2] = True
    m = Model(ndim=3, coords=coords, node_tags=tags_for(8), elem_type=ELEM_HEXA8,
              conn=np.arange(8, dtype=np.int32)[None, :], E=2.08e8, nu=0.30, rho=2000.0, restrained=restrained,
              massform_lumped=False)
    m.add_point_load(5, [0.01, 0.0, 0.0], g["load"])
    m.add_point_load(7, [0.0, 0.01, 0.0], g["load"])
    m.add_point_load(6, [0.0, 0.0, -0.05], g["load"])
    return m


def test_J02_hexa8_consistent_mass_on_device():
    """SVL's J02 case (massform left at SVL's default, CONSISTENT) on the device."""
    with np.load(GOLD / "J02.npz") as z:
        g = {k: z[k] for k in z.files}
    m = j02_model(g)
    rec = [int(d) for d in m.free_dof()[6]]
    ref = oracle.run(m, oracle.NEWMARK, 0.025, 201, rec)
    dm = DeviceModel(m)
    hist = newmark(dm, 0.025).advance(1, 200, rec)
    dm.close()
    err = rel(hist, ref["U"])
    os_u = g["displacement"][:, [3, 1, 2]]  # OpenSees frame (3, 1, 2) = SVL (x, y, z)
    rms = float(np.sqrt(np.mean((hist - os_u) ** 2)) / np.sqrt(np.mean(os_u ** 2)))
    print(f"J02 on the device: vs oracle {err:.2e}, vs OpenSees {rms:.2e} (relative RMS)")
    assert err < 1e-10
    assert rms < 1e-5


def test_newton_j2_consistent_mass():
    from test_gpu_nonlinear import check, j2_block, yield_check
    m, dt, src = j2_block()
    m.massform_lumped = False
    rec = W.dofs_of(m, src) + list(range(0, m.nfree, 17))
    dm, nm = check(m, dt, 100, rec, 1e-8, 30, 5)
    yield_check(nm, m)
    dm.close()


def test_reactions_consistent():
    from test_gpu_reactions import device_history, fixed
    m, dt, nt, src = block_model(4, 3, 3)
    m = consistent(m, 0.5, 1e-3)
    nodes, tot = fixed(m)
    m.add_point_load(int(nodes[4]), [1.0, 0.5, 0.0], np.sin(np.arange(400) * 0.05) * 3e5)
    nt = 40
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, [], reac_tot=tot)["R"]
    dm = DeviceModel(m)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=5000)
    nm.Initialize()
    got = device_history(dm, nm, nodes, nt)
    dm.close()
    err = rel(got, ref)
    print(f"consistent-mass Newmark reactions: rel L2 {err:.2e}")
    assert err < 1e-9


def test_learned_fcg_consistent_mass():
    """The network's operator carries the consistent mass (FP64 element routine); FCG with it
    reproduces the oracle's direct solves and needs fewer iterations than Jacobi-PCG."""
    m, dt, nt = W.block3d_implicit(8, 8, 6, nt=31)
    m.massform_lumped = False
    rec = W.dofs_of(m, int(m.load_node[0])) + list(range(0, m.nfree, 19))
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec)
    dm = DeviceModel(m)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=3000)
    nm.Initialize()
    hist_j = nm.advance(1, nt - 1, rec)
    it_j = nm.iters.mean()
    nm.Initialize()
    pc = LearnedPreconditioner(nm, channels=1, layers=6, mfma=-1)
    assert pc.info()["mfma"] == 0
    pc.use(True)
    hist = nm.advance(1, nt - 1, rec)
    it_l = nm.iters.mean()
    dm.close()
    print(f"consistent mass: Jacobi-PCG {it_j:.1f} vs learned FCG {it_l:.1f} iterations per step; "
          f"rel L2 {rel(hist, ref['U']):.2e} / {rel(hist_j, ref['U']):.2e}")
    assert rel(hist, ref["U"]) < 1e-10 and rel(hist_j, ref["U"]) < 1e-10
    assert it_l < 0.5 * it_j


def test_consistent_mass_refusals():
    """Where the device has no consistent-mass path it says so: CentralDifference (SVL solves
    its non-diagonal Keff by LDL^T) and domain reduction loads."""
    m = hexa_case("axis")
    dm = DeviceModel(m)
    with pytest.raises(SvlgError, match="LUMPED"):
        CentralDifference(dm, 1e-4).Initialize()
    dm.close()
    from svl_amd import drm
    from _meshes import drm_quad_model
    dt, nt = 2e-3, 20
    m, box = drm_quad_model(nt, dt)
    e, n, c = drm.drm_box(m, *box)
    z = np.zeros((nt - 1, m.nfree))
    drm.attach(m, e, n, c, drm.signals_from_history(m, n, nt, z, z, z, 1))
    m.massform_lumped = False
    dm = DeviceModel(m)
    with pytest.raises(SvlgError, match="LUMPED"):
        newmark(dm, dt)
    dm.close()
