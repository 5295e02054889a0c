"""Pins the numpy restatement of PML2DQuad4 and of the mixed-unknown models around it
(oracle/pml2d.py). SVL ships no output for its PML cases (F11 and J12 hold inputs only), so
the pins are mathematical and physical:

* symmetry of the element mass, damping and stiffness matrices;
* without stretching (R = 1: ax = ay = 1, bx = by = 0) on the unit square the stiffness is zero
  and every block has its closed form: rho/36 [4 2 1 2 ...] for the displacement mass, the
  compliance constants times the same pattern for the stress unknowns, and the integrals of
  dN_j/dx N_k (entries of +-1/6, +-1/12) coupling displacements and stresses in the damping;
* with a constant stretching (m = 0) the three matrices scale by ax ay, ax by + ay bx and
  bx by block by block;
* the lin2DQuad4 matrices agree with the C++ oracle's (themselves pinned by closed forms);
* physics: a 1 x 100 soil column driven at the top (SVL's F11 layout: 10 PML elements below
  90 elastic ones, joined by EQUAL constraints, consistent mass, ExtendedNewmarkBeta) follows
  a 400-element column — free of reflections within the window — far more closely than the
  same 90 elements on a fixed base. A sign or coefficient mistake in the layer breaks it.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import pml2d as P
from svl_amd.mesh import ELEM_QUAD4, Model, tags_for

UNIT = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
PAT = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]) / 36.0


def test_pml_matrices_symmetric():
    X = np.array([[0.1, 0.0], [1.2, 0.1], [1.0, 1.3], [-0.1, 0.9]])
    M, C, K = P.pml_matrices(X, 5e7, 0.25, 2000.0, (1.0, 2.0, 5.0, 1e-5, 0.5, 10.0, 0.0, -1.0))
    for A in (M, C, K):
        assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
        assert np.abs(A).max() > 0


def test_closed_forms_without_stretching():
    E, nu, rho = 5e7, 0.25, 2000.0
    lam, mu = P.lame(E, nu)
    M, C, K = P.pml_matrices(UNIT, E, nu, rho, (1.0, 2.0, 5.0, 1.0, 0.0, 0.0, 0.0, 1.0))
    assert np.abs(K).max() == 0.0
    np.testing.assert_allclose(M[0::5, 0::5], rho * PAT, rtol=1e-14)
    np.testing.assert_allclose(M[2::5, 2::5], -(lam + 2 * mu) / (4 * mu * (lam + mu)) * PAT, rtol=1e-14)
    np.testing.assert_allclose(M[4::5, 4::5], -PAT / mu, rtol=1e-14)
    np.testing.assert_allclose(M[2::5, 3::5], lam / (4 * mu * (lam + mu)) * PAT, rtol=1e-14)
    # int dN_j/dx N_k and int dN_j/dy N_k over the unit square, node order (0,0) (1,0) (1,1)
    # (0,1): dN/dx = (-(1-y), 1-y, y, -y) and dN/dy = (-(1-x), -x, x, 1-x), hence entries of
    # +-1/6 where N_j and N_k share the factor in the other coordinate and +-1/12 elsewhere.
    Nx = np.array([[-2, -2, -1, -1], [2, 2, 1, 1], [1, 1, 2, 2], [-1, -1, -2, -2]], float) / 12.0
    Ny = np.array([[-2, -1, -1, -2], [-1, -2, -2, -1], [1, 2, 2, 1], [2, 1, 1, 2]], float) / 12.0
    for (i, j), B in (((0, 2), Nx), ((0, 4), Ny), ((1, 3), Ny), ((1, 4), Nx)):
        np.testing.assert_allclose(C[i::5, j::5], B, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(C[j::5, i::5], B.T, rtol=1e-13, atol=1e-15)
    assert np.abs(C[0::5, 0::5]).max() == 0.0 and np.abs(C[2::5, 2::5]).max() == 0.0


def test_constant_stretching_scales_blocks():
    E, nu, rho = 5e7, 0.25, 2000.0
    lam, mu = P.lame(E, nu)
    p = (1.0, 0.0, 5.0, 1e-3, 0.0, 0.0, 1.0, 1.0)   # m = 0: f = 1 everywhere
    ax, ay, bx, by = P.stretch(0.3, 0.4, rho, mu, lam, p)
    assert ax == ay and bx == by and bx > 0
    M, C, K = P.pml_matrices(UNIT, E, nu, rho, p)
    M0, C0, K0 = P.pml_matrices(UNIT, E, nu, rho, (1.0, 0.0, 5.0, 1.0, 0.0, 0.0, 1.0, 1.0))
    diag = np.zeros((20, 20), bool)
    for a in range(5):
        for b in range(5):
            if (a < 2) == (b < 2):
                diag[a::5, b::5] = True
    np.testing.assert_allclose(M[diag], ax * ay * M0[diag], rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(C[diag], (ax * by + ay * bx) * M0[diag], rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(K[diag], bx * by * M0[diag], rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(C[~diag], ax * C0[~diag], rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(K[~diag], bx * C0[~diag], rtol=1e-13, atol=1e-18)


@pytest.mark.parametrize("lumped", [False, True])
def test_quad_matrices_match_oracle(lumped):
    X = np.array([[0.1, 0.0], [1.2, 0.1], [1.0, 1.3], [-0.1, 0.9]])
    m = Model(ndim=2, coords=X, node_tags=tags_for(4), elem_type=ELEM_QUAD4, conn=np.array([[0, 1, 2, 3]]),
              E=np.array([3e7]), nu=np.array([0.3]), rho=np.array([1800.0]), thickness=0.7)
    m.massform_lumped = lumped
    K, M, _ = oracle.element_matrices(m, 0)
    Kp, Mp = P.quad_matrices(X, 3e7, 0.3, 1800.0, 0.7, lumped)
    np.testing.assert_allclose(Kp, K, rtol=1e-12, atol=1e-12 * np.abs(K).max())
    np.testing.assert_allclose(Mp, M, rtol=1e-12, atol=1e-12 * np.abs(M).max())


def ricker(t, t0=0.3, f0=5.0):
    a = (math.pi * f0 * (t - t0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


def column(n_reg, n_pml, y_base, R=1e-5, L=5.0):
    """1 x n column of unit squares: n_pml PML elements (5 unknowns per node) from y_base, then
    n_reg elastic ones; ux fixed everywhere; the two regions joined by EQUAL constraints on uy.
    Without PML the base is fixed."""
    E, nu, rho = 5e7, 0.25, 2000.0
    coords, ndof, fixed, elems, equal = [], [], [], [], []
    p = (1.0, 2.0, float(L), R, 0.5, y_base + n_pml, 0.0, -1.0)
    if n_pml:
        for j in range(n_pml + 1):
            for i in range(2):
                coords.append([float(i), y_base + j])
                ndof.append(5)
                fixed.append([True, j == 0, False, False, False])
        for j in range(n_pml):
            b = 2 * j
            elems.append(("pml", (b, b + 1, b + 3, b + 2), E, nu, rho, p))
    r0 = len(coords)
    ybot = y_base + n_pml
    for j in range(n_reg + 1):
        for i in range(2):
            coords.append([float(i), ybot + j])
            ndof.append(2)
            fixed.append([True, j == 0 and not n_pml])
    for j in range(n_reg):
        b = r0 + 2 * j
        elems.append(("quad", (b, b + 1, b + 3, b + 2), E, nu, rho, (1.0,)))
    if n_pml:
        top = 2 * n_pml
        equal = [(top, 1, r0, 1, 1.0), (top + 1, 1, r0 + 1, 1, 1.0)]
    m = P.MixedModel(np.array(coords), ndof, fixed, elems, equal, lumped=False)
    return m, len(coords) - 2


def test_pml_column_absorbs():
    dt, nt = 0.01, 200
    s = 1e5 * ricker(np.arange(nt) * dt)
    runs = {}
    for name, (n_reg, n_pml, y0) in {"pml": (90, 10, 0.0), "fixed": (90, 0, 10.0), "long": (400, 0, -300.0)}.items():
        m, top = column(n_reg, n_pml, y0)
        m.loads = [(top, 1, 1.0, s), (top + 1, 1, 1.0, s)]
        runs[name] = m.run_newmark(dt, nt, [(top, 1)])[:, 0]
    ref = runs["long"]
    e_pml = np.linalg.norm(runs["pml"] - ref) / np.linalg.norm(ref)
    e_fix = np.linalg.norm(runs["fixed"] - ref) / np.linalg.norm(ref)
    print(f"PML column error {e_pml:.3e}, fixed-base column error {e_fix:.3e}")
    assert np.all(np.isfinite(runs["pml"]))
    assert e_fix > 0.3
    assert e_pml < 0.1 * e_fix


def test_pml_reflection_coefficient():
    """A wave crossing a continuous PML of depth L twice (fixed end) returns with amplitude
    exp(-2/vp int_0^L b) = R for the profile b0 (d/L)^m, b0 = (m+1) vp ln(1/R) / (2L); the free
    surface doubles it, so the top displacement differs from an unbounded column's by 2R of its
    peak. Checked at R = 0.1 and 0.01 (a 10-element layer, 1 m elements, 10 Hz Ricker pulse);
    below that the mesh's own reflection dominates."""
    dt, nt = 0.005, 200
    s = 1e5 * ricker(np.arange(nt) * dt, t0=0.15, f0=10.0)
    m, top = column(300, 0, -260.0)
    m.loads = [(top, 1, 1.0, s), (top + 1, 1, 1.0, s)]
    ref = m.run_newmark(dt, nt, [(top, 1)])[:, 0]
    for R in (1e-1, 1e-2):
        m, top = column(40, 10, 0.0, R=R, L=10.0)
        m.loads = [(top, 1, 1.0, s), (top + 1, 1, 1.0, s)]
        u = m.run_newmark(dt, nt, [(top, 1)])[:, 0]
        ratio = np.abs(u - ref).max() / np.abs(ref).max() / (2 * R)
        print(f"R={R}: reflected / (2R incident) = {ratio:.3f}")
        assert 0.9 < ratio < 1.25


def test_dense_solve_is_accurate():
    """The oracle's Keff solve against long-double iterative refinement on the F11 column: the
    stress and displacement unknowns differ by eleven orders of magnitude in Keff, so a
    solve that skipped the scaling would miss by ~1e-11 and fail here."""
    m, _ = column(90, 10, 0.0)
    dt = 0.01
    K = m.keff_matrix(dt)
    solve = m.keff_solver(dt)
    b = np.random.default_rng(3).standard_normal(K.shape[0]) * np.sqrt(np.abs(np.diag(K)))
    x = solve(b)
    KL, xl = K.astype(np.longdouble), x.astype(np.longdouble)
    for _ in range(3):
        xl = xl + solve((b.astype(np.longdouble) - KL @ xl).astype(np.float64)).astype(np.longdouble)
    xt = xl.astype(np.float64)
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-14


# ---- 3D (PML3DHexa8, history matrix G, integrated displacement Ubar) -------------------------

CUBE = np.array([[a, b, c] for a, b, c in P.CORNERS3], float) * 0.5 + 0.5   # unit cube, SVL corner order


def _m1(s, t):
    """int_0^1 phi_s phi_t for the 1D hats phi_-(x) = 1 - x, phi_+(x) = x."""
    return 1 / 3 if s == t else 1 / 6


def test_pml3_matrices_symmetric():
    X = CUBE + np.random.default_rng(4).uniform(-0.15, 0.15, CUBE.shape)
    mats = P.pml3_matrices(X, 5e7, 0.25, 2000.0, (2.0, 3.0, 1e-4, 0.2, -0.3, 1.5, 1.0, -1.0, -1.0))
    for A in mats:
        assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
        assert np.abs(A).max() > 0


def test_pml3_closed_forms_without_stretching():
    """R = 1 on the unit cube: K = G = 0; M and C from tensor-product integrals of the hats
    (int N_j N_k = prod of 1/3 or 1/6; int dN_j/dx N_k = s_j/2 * the other two factors)."""
    E, nu, rho = 5e7, 0.25, 2000.0
    lam, mu = P.lame(E, nu)
    M, C, K, G = P.pml3_matrices(CUBE, E, nu, rho, (2.0, 3.0, 1.0, 0, 0, 0, 1, 1, 1))
    assert np.abs(K).max() == 0.0 and np.abs(G).max() == 0.0
    S = P.CORNERS3
    NN = np.array([[_m1(a[0], b[0]) * _m1(a[1], b[1]) * _m1(a[2], b[2]) for b in S] for a in S])
    D = [np.array([[a[d] / 2 * np.prod([_m1(a[e], b[e]) for e in range(3) if e != d]) for b in S] for a in S])
         for d in range(3)]
    np.testing.assert_allclose(M[0::9, 0::9], rho * NN, rtol=1e-13)
    np.testing.assert_allclose(M[3::9, 3::9], -(lam + mu) / mu / (3 * lam + 2 * mu) * NN, rtol=1e-13)
    np.testing.assert_allclose(M[3::9, 5::9], lam / 2 / mu / (3 * lam + 2 * mu) * NN, rtol=1e-13)
    np.testing.assert_allclose(M[7::9, 7::9], -NN / mu, rtol=1e-13)
    for d, pairs in P.COUPLE3.items():
        for pu, ps in pairs:
            np.testing.assert_allclose(C[pu::9, ps::9], D[d], rtol=1e-12, atol=1e-15)
            np.testing.assert_allclose(C[ps::9, pu::9], D[d].T, rtol=1e-12, atol=1e-15)
    assert np.abs(C[0::9, 0::9]).max() == 0.0


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_stretching_laplace_identity(dim):
    """With the same stretching s(p) = a + b/p along every axis (m = 0), the coordinate
    stretch multiplies the mass-like operator by s^dim and each derivative by 1/s. In the
    Laplace domain the element operator p^2 M + p C + K + G/p therefore equals
    p^2 s^2 M0 + p s C0 in 2D and p^2 s^3 M0 + p s^2 C0 in 3D, M0 and C0 the unstretched (R = 1)
    mass and coupling. Checked at complex p; this fixes every stretching coefficient of M, C,
    K and G (each is a binomial term of these products)."""
    E, nu, rho = 5e7, 0.25, 2000.0
    lam, mu = P.lame(E, nu)
    if dim == 2:
        X = UNIT + np.array([[0.0, 0.0], [0.1, 0.05], [0.05, 0.1], [-0.05, 0.0]])
        M, C, K = P.pml_matrices(X, E, nu, rho, (1.0, 0.0, 4.0, 1e-3, 0.0, 0.0, 1.0, 1.0))
        G = np.zeros_like(M)
        M0, C0, _ = P.pml_matrices(X, E, nu, rho, (1.0, 0.0, 4.0, 1.0, 0.0, 0.0, 1.0, 1.0))
        a, _, b, _ = P.stretch(0.0, 0.0, rho, mu, lam, (1.0, 0.0, 4.0, 1e-3, 0.0, 0.0, 1.0, 1.0))
    else:
        X = CUBE + np.random.default_rng(5).uniform(-0.1, 0.1, CUBE.shape)
        p9 = (0.0, 4.0, 1e-3, 0.0, 0.0, 0.0, 1.0, -1.0, 1.0)
        M, C, K, G = P.pml3_matrices(X, E, nu, rho, p9)
        M0, C0, _, _ = P.pml3_matrices(X, E, nu, rho, p9[:2] + (1.0,) + p9[3:])
        a, _, _, b, _, _ = P.stretch3(0.0, 0.0, 0.0, rho, mu, lam, p9)
    assert b > 0 and a > 1
    for pz in (0.7 + 2.0j, 3.0 - 1.5j, 40.0 + 90.0j):
        s = a + b / pz
        Z = pz ** 2 * M + pz * C + K + G / pz
        ref = (pz ** 2 * s ** 2 * M0 + pz * s * C0) if dim == 2 else (pz ** 2 * s ** 3 * M0 + pz * s ** 2 * C0)
        assert np.abs(Z - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("lumped", [False, True])
def test_hexa_matrices_match_oracle(lumped):
    X = CUBE * [1.2, 0.9, 1.1] + np.random.default_rng(6).uniform(-0.1, 0.1, CUBE.shape)
    m = Model(ndim=3, coords=X, node_tags=tags_for(8), elem_type=1, conn=np.arange(8)[None, :],
              E=np.array([3e7]), nu=np.array([0.3]), rho=np.array([1800.0]))
    m.massform_lumped = lumped
    K, M, _ = oracle.element_matrices(m, 0)
    Kp, Mp = P.hexa_matrices(X, 3e7, 0.3, 1800.0, lumped)
    np.testing.assert_allclose(Kp, K, rtol=1e-12, atol=1e-12 * np.abs(K).max())
    np.testing.assert_allclose(Mp, M, rtol=1e-12, atol=1e-12 * np.abs(M).max())


def test_rod3d_reflection_coefficient():
    """SVL's J12 layout (PML3DHexa8 under lin3DHexa8, EQUAL ties on uz): the top response of a
    40-element rod on a 10-element layer differs from an unbounded rod's by about 2R."""
    from _pml_meshes import to_oracle
    from svl_amd.pml import soil_rod3d
    dt, nt = 0.005, 200
    s = ricker(np.arange(nt) * dt, t0=0.15, f0=10.0)
    ref_m = to_oracle(soil_rod3d(300, 1, s, R=1.0))     # a 1-element layer with R = 1 is a plain free end far below
    top = len(ref_m.ndof) - 4
    ref = ref_m.run_newmark(dt, nt, [(top, 2)])[:, 0]
    for R in (1e-1, 1e-2):
        m = to_oracle(soil_rod3d(40, 10, s, R=R))
        u = m.run_newmark(dt, nt, [(len(m.ndof) - 4, 2)])[:, 0]
        ratio = np.abs(u - ref).max() / np.abs(ref).max() / (2 * R)
        print(f"3D rod R={R}: reflected / (2R incident) = {ratio:.3f}")
        assert 0.9 < ratio < 1.25


def test_newmark_with_history_matrix():
    """A 3D half-space with PML on five faces (corner elements stretched in three directions,
    so G != 0): every step of the oracle satisfies the equation of motion at t_k,
    T^T (M A + C V + K U + G Ubar - F) = 0, and Ubar grows by the exact integral of the
    piecewise-quadratic displacement of constant average acceleration, written here as the
    Hermite rule dt/2 (U_k-1 + U_k) + dt^2/12 (V_k-1 - V_k)."""
    from _pml_meshes import to_oracle
    from svl_amd.pml import halfspace_box3d
    dt, nt = 0.005, 40
    mm = to_oracle(halfspace_box3d(2, 2, 2, 1, load_series=ricker(np.arange(nt) * dt, t0=0.05, f0=20.0)))
    off, T, M, C, K = mm.assemble()
    G = mm._G
    assert np.abs(G).max() > 0
    U, V, A, Ub, F = mm.run_newmark(dt, nt, [], full=True)
    scale = np.abs(T.T @ F.T).max()
    for k in range(U.shape[0]):
        r = T.T @ (M @ A[k] + C @ V[k] + K @ U[k] + G @ Ub[k] - F[k])
        assert np.abs(r).max() <= 1e-8 * scale
    U0 = np.vstack([np.zeros(U.shape[1]), U[:-1]])
    V0 = np.vstack([np.zeros(V.shape[1]), V[:-1]])
    Ub0 = np.vstack([np.zeros(Ub.shape[1]), Ub[:-1]])
    herm = Ub0 + dt / 2 * (U0 + U) + dt * dt / 12 * (V0 - V)
    np.testing.assert_allclose(Ub, herm, rtol=0, atol=1e-12 * np.abs(Ub).max())
    assert np.abs(Ub).max() > 0
