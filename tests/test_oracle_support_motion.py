"""Oracle: support motion (Supports + POINTLOAD_SUPPORT_MOTION) pinned by what the mathematics
fixes, without the oracle's own support-motion code on the reference side.

SVL's step k adds a_k (the increment of sample k, NewmarkBeta: k - 1, Assembler.cpp:493-532)
to the restrained DOFs, subtracts T^T K a_k from the effective force, and its element strains
are updated before a_k enters U (Algorithm.cpp:18-56), so the internal force of step k sees the
support displacement G_(k-2) instead of G_(k-1) (G_j = a_1 + ... + a_j, 0 for j <= 0).

When every restrained DOF of direction x carries the same motion and nothing ties the body to
the ground (no dashpots; mass-proportional damping only, lumped mass), K 1_x = 0 and the free
DOFs follow a fixed-base run of the same model under point loads on the free x DOFs:
  CentralDifference:  U_k = u_k + G_(k-1) 1_x,  F'_k = F_k - M 1 (G_(k-1) - 2 G_(k-2) + G_(k-3))/dt^2
                      - am M 1 (G_(k-1) - G_(k-3)) / 2dt;
  ExtendedNewmarkBeta / NewmarkBeta (gamma 1/2, beta 1/4): U_k = u_k + H_k 1_x with
                      H_k = G_(k-2) + a_k, dH_k = H_k - H_(k-1), P_k = 2/dt dH_k - P_(k-1),
                      Q_k = 4/dt^2 dH_k - 4/dt P_(k-1) - Q_(k-1), F'_k = F_k - M 1 (Q_k + am P_k)
(substituting U = u + H 1 into the step equations: the K terms cancel through K 1_x = 0 and the
velocity and acceleration shifts P, Q carry the rest). The fixed-base runs use the oracle's point
loads, pinned by SVL's own fixtures. A dropped K a_k term, a sign error, no lag, or the wrong
sample index fails these identities; the restrained history must equal the closed forms
g(k) - g(0) (sample k) and g(k - 1) (NewmarkBeta).
"""
import numpy as np
import pytest

import oracle
from _meshes import support_model


def applied(g, factor, nt, offset):
    """a_k, k = 0..nt-1 (a_0 = 0: no step 0), from the closed-form displacement g~(k)."""
    n = len(g)

    def val(j):
        return g[j] if j < n else g[0]

    gt = np.array([0.0] + [factor * (val(k - offset) if k - offset >= 0 else 0.0) for k in range(1, nt)])
    base = factor * (val(0) if offset == 0 else 0.0)
    # G_k = g~(k) - g~(0 step): offset 0: g(k) - g(0); offset 1: g(k-1) (with g(-1) = 0).
    G = np.where(np.arange(nt) >= 1, gt - base, 0.0)
    a = np.diff(np.concatenate([[0.0], G]))
    return a, G


def shifted(G, j):
    return G[j] if j >= 1 else 0.0


def x_dofs(m):
    fd = m.free_dof()
    return fd[:, 0][fd[:, 0] >= 0]


def fixed_base_copy(m):
    import copy
    m0 = copy.deepcopy(m)
    m0.sm_node = np.zeros(0, np.int32)
    m0.sm_dir = np.zeros(0, np.int32)
    m0.sm_factor = np.zeros(0)
    m0.sm_series = []
    return m0


def add_loads(m0, m, loads):
    """One point load row per node with a free x DOF: series loads[:, free x index]."""
    fd = m.free_dof()
    for n in range(m.nnodes):
        f = fd[n, 0]
        if f < 0:
            continue
        d = np.zeros(m.ndim)
        d[0] = 1.0
        m0.add_point_load(n, d, loads[:, f])


def restrained_x(m, bnd):
    return [int(n) for n in bnd["bottom"]]


@pytest.mark.parametrize("ndim,am", [(2, 0.0), (2, 0.8), (3, 0.5)])
def test_central_difference_relative_coordinates(ndim, am):
    dt, nt = 2e-4, 500
    m, g, bnd = support_model(ndim, nt, dt, am=am, length=350, factor=0.8)
    nf = m.nfree
    rec = list(range(nf))
    base = restrained_x(m, bnd)
    rec_all = rec + [-1 - (n * ndim) for n in base]
    out = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec_all)["U"]
    a, G = applied(g, 0.8, nt, 0)
    Mf, _ = oracle.diagonals(m)
    loads = np.zeros((nt, nf))
    xd = x_dofs(m)
    for k in range(1, nt):
        acc = (shifted(G, k - 1) - 2.0 * shifted(G, k - 2) + shifted(G, k - 3)) / dt / dt
        vel = (shifted(G, k - 1) - shifted(G, k - 3)) / (2.0 * dt)
        loads[k, xd] = -Mf[xd] * (acc + am * vel)
    m0 = fixed_base_copy(m)
    add_loads(m0, m, loads)
    u = oracle.run(m0, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)["U"]
    want = u.copy()
    for k in range(1, nt):
        want[k - 1, xd] += shifted(G, k - 1)
    err = np.linalg.norm(out[:, :nf] - want) / np.linalg.norm(want)
    print(f"CD {ndim}D am={am}: free DOFs vs fixed-base relative run {err:.2e}")
    assert err < 1e-10
    # Restrained DOFs: g(k) - g(0) times the factor.
    assert np.allclose(out[:, nf:], G[1:, None], rtol=0, atol=1e-15)


@pytest.mark.parametrize("integrator,offset,ndim,am", [
    (oracle.EXTENDED_NEWMARK, 0, 2, 0.0),
    (oracle.EXTENDED_NEWMARK, 0, 3, 0.6),
    (oracle.NEWMARK, 1, 2, 0.4),
])
def test_newmark_relative_coordinates(integrator, offset, ndim, am):
    dt, nt = 1e-3, 160
    m, g, bnd = support_model(ndim, nt, dt, am=am, length=110, factor=1.3)
    nf = m.nfree
    rec = list(range(nf))
    base = restrained_x(m, bnd)
    rec_all = rec + [-1 - (n * ndim) for n in base]
    out = oracle.run(m, integrator, dt, nt, rec_all)["U"]
    a, G = applied(g, 1.3, nt, offset)
    H = np.zeros(nt)
    P = np.zeros(nt)
    Q = np.zeros(nt)
    for k in range(1, nt):
        H[k] = shifted(G, k - 2) + a[k]
        dH = H[k] - H[k - 1]
        P[k] = 2.0 / dt * dH - P[k - 1]
        Q[k] = 4.0 / dt / dt * dH - 4.0 / dt * P[k - 1] - Q[k - 1]
    Mf, _ = oracle.diagonals(m)
    xd = x_dofs(m)
    loads = np.zeros((nt, nf))
    for k in range(1, nt):
        loads[k, xd] = -Mf[xd] * (Q[k] + am * P[k])
    m0 = fixed_base_copy(m)
    add_loads(m0, m, loads)
    u = oracle.run(m0, integrator, dt, nt, rec)["U"]
    want = u.copy()
    for k in range(1, nt):
        want[k - 1, xd] += H[k]
    err = np.linalg.norm(out[:, :nf] - want) / np.linalg.norm(want)
    print(f"Newmark ({integrator}) {ndim}D am={am}: free DOFs vs fixed-base relative run {err:.2e}")
    assert err < 1e-10
    assert np.allclose(out[:, nf:], G[1:, None], rtol=0, atol=1e-15)


def test_restrained_history_closed_forms():
    """U at the support DOFs: factor (g(k) - g(0)) for sample k, factor g(k - 1) for NewmarkBeta
    (its first step applies g(0) itself); past the series SVL reads g(0) again."""
    dt, nt = 1e-3, 60
    m, g, bnd = support_model(2, nt, dt, length=40, factor=0.5, load=False)
    n0 = int(bnd["bottom"][2])
    val = lambda j: g[j] if j < len(g) else g[0]  # noqa: E731
    for integ, want in ((oracle.CENTRAL_DIFFERENCE, [0.5 * (val(k) - val(0)) for k in range(1, nt)]),
                        (oracle.EXTENDED_NEWMARK, [0.5 * (val(k) - val(0)) for k in range(1, nt)]),
                        (oracle.NEWMARK, [0.5 * val(k - 1) for k in range(1, nt)])):
        dt_i = 2e-4 if integ == oracle.CENTRAL_DIFFERENCE else dt
        got = oracle.run(m, integ, dt_i, nt, [-1 - n0 * 2])["U"][:, 0]
        assert np.allclose(got, want, rtol=0, atol=1e-15), integ
    # Two loads on one DOF add their factors' increments.
    m2, g2, bnd2 = support_model(2, nt, dt, length=40, factor=0.5, load=False)
    for n in bnd2["bottom"]:
        m2.add_support_motion(int(n), 0, g2, 0.25)
    got = oracle.run(m2, oracle.EXTENDED_NEWMARK, dt, nt, [-1 - n0 * 2])["U"][:, 0]
    assert np.allclose(got, [0.75 * (val(k) - val(0)) for k in range(1, nt)], rtol=1e-14, atol=1e-18)


def test_newton_with_support_motion_is_refused():
    m, g, bnd = support_model(2, 20, 1e-3)
    with pytest.raises(RuntimeError, match="Newton"):
        oracle.run(m, oracle.NEWMARK, 1e-3, 20, [0], algorithm=oracle.NEWTON)
