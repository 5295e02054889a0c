"""The FP64 oracle against closed forms and invariants (CPU).

Each check uses a derivation independent of the oracle's code path, chosen so that a
plausible slip (a dropped term, a wrong sign or index, a transposed operand) fails:
  * element stiffness of the unit square / unit cube from exact tensor-product
    integrals of the bilinear / trilinear shape-function derivatives (no Jacobian, no
    Gauss points);
  * rotation covariance, symmetry, the rigid-body null space, and the patch test on
    distorted elements;
  * lumped mass = rho V / nen per node for rectangles and bricks;
  * the exact discrete solutions of SVL's central-difference recurrence (undamped and
    damped) and of the trapezoidal Newmark rule for a one-DOF model;
  * the Plain DOF numbering.
"""
import itertools

import numpy as np
import pytest

import oracle
from svl_amd.mesh import ELEM_HEXA8, ELEM_QUAD4, Model, make_domain_area, make_domain_volume, tags_for

E, NU, RHO = 1.7e8, 0.27, 1900.0
LAM = E * NU / (1 - 2 * NU) / (1 + NU)
MU = E / (2 * (1 + NU))


def quad(coords, **kw):
    return Model(ndim=2, coords=np.asarray(coords, float), node_tags=tags_for(4), elem_type=ELEM_QUAD4,
                 conn=np.array([[0, 1, 2, 3]]), E=E, nu=NU, rho=RHO, **kw)


def hexa(coords, **kw):
    return Model(ndim=3, coords=np.asarray(coords, float), node_tags=tags_for(8), elem_type=ELEM_HEXA8,
                 conn=np.arange(8, dtype=np.int32)[None, :], E=E, nu=NU, rho=RHO, **kw)


UNIT_SQUARE = [[0, 0], [1, 0], [1, 1], [0, 1]]
UNIT_CUBE = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]


def exact_stiffness(dim):
    """K of the unit square / cube from int dN_a/dx_p dN_b/dx_q over [0,1]^dim.

    With N_a = prod_i phi_{s_ai}(x_i), phi_0 = 1 - x, phi_1 = x:
      int phi_p phi_q = 1/3 (p == q) else 1/6,  int phi_p' phi_q = s_p / 2,
      int phi_p' phi_q' = s_p s_q  with s_0 = -1, s_1 = +1.
    K_(a i),(b j) = sum_pq int dN_a/dx_p dN_b/dx_q C_{p i q j} with isotropic C.
    """
    corners = np.array(UNIT_SQUARE if dim == 2 else UNIT_CUBE, dtype=int)
    n = corners.shape[0]
    sgn = 2 * corners - 1

    def integral(a, b, p, q):
        val = 1.0
        for ax in range(dim):
            da, db = ax == p, ax == q
            sa, sb = sgn[a, ax], sgn[b, ax]
            if da and db:
                val *= sa * sb
            elif da:
                val *= sa / 2.0
            elif db:
                val *= sb / 2.0
            else:
                val *= 1 / 3 if corners[a, ax] == corners[b, ax] else 1 / 6
        return val

    def C(p, i, q, j):
        return LAM * (p == i) * (q == j) + MU * ((p == q) * (i == j) + (p == j) * (i == q))

    K = np.zeros((n * dim, n * dim))
    for a, b in itertools.product(range(n), repeat=2):
        for i, j in itertools.product(range(dim), repeat=2):
            K[a * dim + i, b * dim + j] = sum(integral(a, b, p, q) * C(p, i, q, j)
                                              for p in range(dim) for q in range(dim))
    return K


@pytest.mark.parametrize("dim", [2, 3])
def test_stiffness_matches_exact_integrals(dim):
    m = quad(UNIT_SQUARE) if dim == 2 else hexa(UNIT_CUBE)
    K, M, _ = oracle.element_matrices(m)
    Kx = exact_stiffness(dim)
    assert np.allclose(K, Kx, rtol=0, atol=1e-14 * np.abs(Kx).max() * 10)


def rotation3(a, b, c):
    def R(axis, t):
        c_, s_ = np.cos(t), np.sin(t)
        M = np.eye(3)
        i, j = [k for k in range(3) if k != axis]
        M[i, i], M[i, j], M[j, i], M[j, j] = c_, -s_, s_, c_
        return M
    return R(2, a) @ R(1, b) @ R(0, c)


def test_rotation_covariance_hexa():
    """K(R x) = (I (x) R) K(x) (I (x) R)^T for an isotropic material."""
    rng = np.random.default_rng(3)
    X = np.array(UNIT_CUBE, float) * np.array([2.0, 1.5, 0.7]) + rng.uniform(-0.15, 0.15, (8, 3))
    R = rotation3(0.4, -0.7, 1.1)
    K1, _, _ = oracle.element_matrices(hexa(X))
    K2, _, _ = oracle.element_matrices(hexa(X @ R.T))
    T = np.kron(np.eye(8), R)
    assert np.allclose(K2, T @ K1 @ T.T, atol=1e-9 * np.abs(K1).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_distorted_element_invariants(dim):
    rng = np.random.default_rng(11)
    base = np.array(UNIT_SQUARE if dim == 2 else UNIT_CUBE, float)
    X = base * 2.0 + rng.uniform(-0.3, 0.3, base.shape)
    m = quad(X) if dim == 2 else hexa(X)
    K, M, _ = oracle.element_matrices(m)
    assert np.allclose(K, K.T, atol=1e-12 * np.abs(K).max())
    w = np.linalg.eigvalsh(K)
    nrigid = 3 if dim == 2 else 6
    assert np.all(w[:nrigid] < 1e-9 * w[-1]) and w[nrigid] > 1e-6 * w[-1]
    # Rigid translations and infinitesimal rotations are in the null space.
    modes = []
    for d in range(dim):
        t = np.zeros((base.shape[0], dim)); t[:, d] = 1.0; modes.append(t.ravel())
    pairs = [(0, 1)] if dim == 2 else [(0, 1), (1, 2), (0, 2)]
    for i, j in pairs:
        r = np.zeros((base.shape[0], dim)); r[:, i] = -X[:, j]; r[:, j] = X[:, i]; modes.append(r.ravel())
    for r in modes:
        assert np.linalg.norm(K @ r) < 1e-9 * np.abs(K).max() * np.linalg.norm(r)
    # Lumped mass: positive diagonal that sums to rho * volume per component.
    assert np.allclose(M, np.diag(np.diag(M)))


@pytest.mark.parametrize("dim", [2, 3])
def test_patch_test_distorted_mesh(dim):
    """A linear displacement field gives zero net internal force at interior nodes."""
    rng = np.random.default_rng(5)
    if dim == 2:
        coords, conn, bnd = make_domain_area(4, 3, (0, 0), (4, 0), (0, 3))
        etype = ELEM_QUAD4
    else:
        coords, conn, bnd = make_domain_volume(3, 3, 2, (0, 0, 0), (3, 0, 0), (0, 3, 0), (0, 0, 2))
        etype = ELEM_HEXA8
    boundary = np.unique(np.concatenate(list(bnd.values())))
    interior = np.setdiff1d(np.arange(coords.shape[0]), boundary)
    coords = coords + np.isin(np.arange(coords.shape[0]), interior)[:, None] * rng.uniform(-0.25, 0.25, coords.shape)
    m = Model(ndim=dim, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=etype, conn=conn,
              E=E, nu=NU, rho=RHO)
    A = rng.standard_normal((dim, dim)) * 1e-3
    U = (coords @ A.T).ravel()
    F = oracle.internal_force(m, U, ftol=0.0).reshape(-1, dim)
    fmax = np.abs(F).max()
    assert fmax > 0
    assert np.abs(F[interior]).max() < 1e-9 * fmax
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * fmax


def test_lumped_mass_of_bricks_and_rectangles():
    m = hexa(np.array(UNIT_CUBE, float) * [2.0, 3.0, 0.5])
    _, M, _ = oracle.element_matrices(m)
    assert np.allclose(np.diag(M), RHO * 3.0 / 8)
    q = quad(np.array(UNIT_SQUARE, float) * [1.5, 0.4], thickness=0.3)
    _, M, _ = oracle.element_matrices(q)
    assert np.allclose(np.diag(M), RHO * 1.5 * 0.4 * 0.3 / 4)


def sdof_model(c=0.0, F0=1.0e4, nt=400):
    """One free DOF (x of the corner (1,1) of a unit square), optional dashpot on it."""
    restrained = np.ones((4, 2), bool)
    restrained[2, 0] = False
    kw = {}
    if c:
        kw = dict(dash_node=np.array([2]), dash_dir=np.array([0]), dash_eta=np.array([c]))
    m = quad(UNIT_SQUARE, restrained=restrained, **kw)
    m.add_point_load(2, [1.0, 0.0], np.full(nt, F0))
    return m, RHO / 4.0, (LAM + 2 * MU + MU) / 3.0


def test_central_difference_sdof_exact_discrete_solution():
    m, mass, k = sdof_model()
    w = np.sqrt(k / mass)
    dt = 0.3 * 2.0 / w
    nt = 400
    r = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, [0])
    Om = w * dt
    th = 2.0 * np.arcsin(Om / 2.0)
    n = np.arange(1, nt)
    up = 1.0e4 / k
    exact = up * (1.0 - np.cos(th * n) + np.tan(th / 2.0) * np.sin(th * n))
    assert np.max(np.abs(r["U"][:, 0] - exact)) < 1e-12 * up * 10


def test_central_difference_damped_sdof():
    m0, mass, k = sdof_model()
    w = np.sqrt(k / mass)
    c = 0.1 * 2 * mass * w
    m, _, _ = sdof_model(c=c)
    dt = 0.25 * 2.0 / w
    nt = 400
    r = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, [0])
    a, b, d = mass / dt ** 2 + c / (2 * dt), 2 * mass / dt ** 2 - k, mass / dt ** 2 - c / (2 * dt)
    rho = np.sqrt(d / a)
    th = np.arccos(b / (2 * np.sqrt(a * d)))
    up = 1.0e4 / k
    A = -up
    B = up * (rho - np.cos(th)) / np.sin(th)
    n = np.arange(1, nt)
    exact = up + rho ** n * (A * np.cos(th * n) + B * np.sin(th * n))
    assert np.max(np.abs(r["U"][:, 0] - exact)) < 1e-11 * up


def test_newmark_sdof_exact_discrete_solution():
    m, mass, k = sdof_model()
    w = np.sqrt(k / mass)
    dt = 1.7 / w
    nt = 300
    r = oracle.run(m, oracle.NEWMARK, dt, nt, [0], want_va=True)
    up = 1.0e4 / k
    u1 = 1.0e4 / (k + 4 * mass / dt ** 2)
    v1 = 2.0 / dt * u1
    th = 2.0 * np.arctan(w * dt / 2.0)
    n = np.arange(1, nt)
    wv, zv = u1 - up, v1 / w
    exact_u = up + wv * np.cos(th * (n - 1)) + zv * np.sin(th * (n - 1))
    exact_v = w * (-wv * np.sin(th * (n - 1)) + zv * np.cos(th * (n - 1)))
    assert np.max(np.abs(r["U"][:, 0] - exact_u)) < 1e-11 * up
    assert np.max(np.abs(r["V"][:, 0] - exact_v)) < 1e-10 * up * w


def test_plain_numbering():
    """Plain scheme: nodes in ascending tag, components in order, restrained skipped
    (01-Pre_Process/Core/Numberer.py:207-228)."""
    coords, conn, bnd = make_domain_area(3, 2, (0, 0), (3, 0), (0, 2))
    restrained = np.zeros(coords.shape, bool)
    restrained[bnd["bottom"], 1] = True
    restrained[0, :] = True
    m = Model(ndim=2, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_QUAD4, conn=conn,
              E=E, nu=NU, rho=RHO, restrained=restrained)
    f = m.free_dof()
    expect, k = [], 0
    for n in range(coords.shape[0]):
        row = []
        for c in range(2):
            if restrained[n, c]:
                row.append(-1)
            else:
                row.append(k)
                k += 1
        expect.append(row)
    assert np.array_equal(f, np.array(expect))
    assert m.nfree == k
