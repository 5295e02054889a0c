"""Pins the oracle's reactions (Integrator::ComputeReactionForce: R = Fdyn - Fext at the fixed
nodes, Fdyn = sum over elements with a fixed node of Fint_e + C_e V_e + M_e A_e, dashpots C_e V_e).

* Balance (Newmark): the discrete equation M A + C V + K U = Fext holds at every free DOF, the
  element elastic forces and the dashpot forces are self-equilibrated and K's Rayleigh part
  too, so per direction the reactions of all fixed nodes sum to
  sum_free (M (A + am V)) - sum Fext — a dropped inertia or mass-damping term, a wrong sign or
  a load counted twice breaks it;
* dashpot partner nodes carry -eta v of the mesh node (sign and coefficient of the dashpot);
* per fixed DOF, R equals K_e u + C_e v + M_e a assembled in numpy from the element matrices
  (themselves pinned by the closed-form tests) minus the load, with lumped and consistent
  mass: catches an index or sign mistake of the assembly, including the elastic part the
  balance cannot see and the inertia of free neighbours that only a consistent mass couples.
"""
import numpy as np
import pytest

import oracle

from _meshes import block_model


def model():
    m, dt, nt, src = block_model(4, 3, 3)
    m.ray_am = np.full(m.nelem, 0.5)
    m.ray_ak = np.full(m.nelem, 1e-3)
    m.add_point_load(int(np.nonzero(m.restrained.any(1))[0][5]), [1.0, 0.5, 0.0],
                     np.sin(np.arange(400) * 0.05) * 3e5)  # a load on a fixed node enters Fext there
    return m, dt


def partners(m):
    order = []
    for n in m.dash_node:
        if int(n) not in order:
            order.append(int(n))
    return {n: m.nnodes + i for i, n in enumerate(order)}


def fixed_total_dofs(m):
    nd = m.ndim
    fixed_nodes = np.nonzero(m.restrained.any(1))[0]
    dofs = [int(n) * nd + c for n in fixed_nodes for c in range(nd)]
    part = partners(m)
    dofs += [p * nd + c for p in part.values() for c in range(nd)]
    return dofs, fixed_nodes, part


def fext(m, k):
    F = np.zeros((m.nnodes, m.ndim))
    for l in range(m.load_node.size):
        F[m.load_node[l]] += m.load_factor[l] * (m.load_series[l, k] * m.load_dir[l])
    return F


@pytest.fixture(scope="module", params=[True, False], ids=["lumped", "consistent"])
def run(request):
    m, dt = model()
    m.massform_lumped = request.param
    nt = 30
    tot, fixed_nodes, part = fixed_total_dofs(m)
    r = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, list(range(m.nfree)), want_va=True, reac_tot=tot)
    return m, dt, nt, tot, fixed_nodes, part, r


def test_total_reaction_balances_inertia_and_load(run):
    m, dt, nt, tot, fixed_nodes, part, r = run
    # With the consistent mass the balance holds with the lumped diagonal: summing M (A + am V)
    # over all rows of a component (restrained DOFs at rest) weighs each DOF by its column sum,
    # which is its row-sum lumped entry.
    import copy
    ml = copy.copy(m)
    ml.massform_lumped = True
    Md, _ = oracle.diagonals(ml)
    fd = m.free_dof()
    comp = np.zeros(m.nfree, int)
    for n in range(m.nnodes):
        for c in range(m.ndim):
            if fd[n, c] >= 0:
                comp[fd[n, c]] = c
    am = 0.5
    for k in range(1, nt):
        row = k - 1
        F = fext(m, k)
        for c in range(m.ndim):
            lhs = sum(r["R"][row, i] for i, t in enumerate(tot) if t % m.ndim == c)
            sel = comp == c
            rhs = np.sum(Md[sel] * (r["A"][row, sel] + am * r["V"][row, sel])) - F[:, c].sum()
            assert abs(lhs - rhs) <= 1e-8 * (np.abs(r["R"][row]).sum() + abs(rhs) + 1.0)


def test_dashpot_partner_reaction(run):
    m, dt, nt, tot, fixed_nodes, part, r = run
    fd = m.free_dof()
    for n, p in list(part.items())[:6]:
        for c in range(m.ndim):
            eta = sum(m.dash_eta[i] for i in range(m.dash_node.size) if m.dash_node[i] == n and m.dash_dir[i] == c)
            i = tot.index(p * m.ndim + c)
            v = r["V"][:, fd[n, c]] if fd[n, c] >= 0 else 0.0
            assert np.allclose(r["R"][:, i], -eta * v, rtol=1e-12, atol=1e-9)


def test_fixed_node_reactions_against_element_matrices(run):
    m, dt, nt, tot, fixed_nodes, part, r = run
    nd, nen = m.ndim, m.nen
    fd = m.free_dof()
    row = nt - 2
    U, V, A = (np.zeros((m.nnodes, nd)) for _ in range(3))
    for n in range(m.nnodes):
        for c in range(nd):
            if fd[n, c] >= 0:
                U[n, c], V[n, c], A[n, c] = r["U"][row, fd[n, c]], r["V"][row, fd[n, c]], r["A"][row, fd[n, c]]
    R = np.zeros((m.nnodes, nd))
    fixed = set(int(n) for n in fixed_nodes)
    for e in range(m.nelem):
        nodes = m.conn[e]
        if not any(int(n) in fixed for n in nodes):
            continue
        K, M, Cm = oracle.element_matrices(m, e)
        f = K @ U[nodes].reshape(-1) + Cm @ V[nodes].reshape(-1) + M @ A[nodes].reshape(-1)
        np.add.at(R, nodes, f.reshape(nen, nd))
    for i in range(m.dash_node.size):
        n, c = int(m.dash_node[i]), int(m.dash_dir[i])
        if n in fixed:
            R[n, c] += m.dash_eta[i] * V[n, c]
    R -= fext(m, row + 1)
    for n in fixed_nodes:
        for c in range(nd):
            got = r["R"][row, tot.index(int(n) * nd + c)]
            assert abs(got - R[n, c]) <= 1e-9 * (np.abs(R).max() + 1.0)
