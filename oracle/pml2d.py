"""PML2DQuad4 and the mixed-unknown 2D models around it: a numpy restatement (test
infrastructure only; no product code calls it).

SVL couples a perfectly matched layer of PML2DQuad4 elements to lin2DQuad4 elements through
EQUAL constraints between coincident nodes. A PML node carries five unknowns (ux, uy and the
stress-like components xx, yy, xy); a regular node two. This module restates, for small
dense models:

* PML2DQuad4 mass, damping and stiffness matrices (20 x 20, node-major, five unknowns per
  node) with the stretching functions of PML2DQuad4::ComputePMLStretchingFactors
  (RP/04-Elements/06-Quadrilateral/PML2DQuad4.cpp:233-466, 655-693; parameters th, m, L, R,
  x0, y0, nx, ny in the order of the constructor, :11-12), 2 x 2 Gauss points with SVL's
  abscissa; the 2D element has no history matrix (ComputePMLMatrix is empty, :467-471);
* lin2DQuad4 plane-strain stiffness and consistent / row-sum lumped mass (8 x 8);
* assembly in element order, the total-to-free map with EQUAL constraints (a slave unknown
  takes the master's free number times the factor, RP/06-Mesh/Mesh.cpp:328-383) and
  ExtendedNewmarkBeta with the Linear algorithm and G = 0 (ExtendedNewmarkBeta.cpp:123-154,
  208-245), solved densely after a symmetric diagonal scaling (keff_solver).

Pins: tests/test_oracle_pml.py (symmetry, closed forms without stretching, constant-stretching
scaling, the lin2DQuad4 matrices against the C++ oracle, a soil column whose PML-truncated
response follows a much longer column while a fixed truncation does not, and the reflection
coefficient 2R at the free surface that fixes the damping profile b0).

Parity unpinned: the evanescent stretch coefficient a0 beyond its R = 1 and m = 0 cases. It
does not enter the reflection of a propagating wave in the continuous layer, and SVL ships no
PML output to compare against.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

G = 0.577350269189626  # SVL's 2-point Gauss abscissa (GaussQuadrature.cpp:303-316)
GP = [(-G, -G), (G, -G), (-G, G), (G, G)]  # x fastest, weights 1


def shape(r, s):
    return np.array([(1 - r) * (1 - s), (1 + r) * (1 - s), (1 + r) * (1 + s), (1 - r) * (1 + s)]) / 4.0


def dshape(r, s):
    """[dN/dr, dN/ds] rows of the four corners."""
    return np.array([[-(1 - s), (1 - s), (1 + s), -(1 + s)],
                     [-(1 - r), -(1 + r), (1 + r), (1 - r)]]) / 4.0


def jac(X, r, s):
    return dshape(r, s) @ X  # J[k][j] = d x_j / d xi_k


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def stretch(x, y, rho, mu, lam, p):
    """(ax, ay, bx, by) at (x, y); p = (th, m, L, R, x0, y0, nx, ny) (PML2DQuad4.cpp:655-693)."""
    th, m, L, R, x0, y0, nx, ny = p
    vp = math.sqrt((lam + 2 * mu) / rho)
    b_pml = L / 10.0
    a0 = (m + 1) * b_pml / 2 / L * math.log(1 / R)
    b0 = (m + 1) * vp / 2 / L * math.log(1 / R)
    fx = ((x - x0) * nx / L) ** m
    fy = ((y - y0) * ny / L) ** m
    return 1 + a0 * fx, 1 + a0 * fy, b0 * fx, b0 * fy


def pml_matrices(X, E, nu, rho, p):
    """M, C, K (20 x 20) of PML2DQuad4 (PML2DQuad4.cpp:233-466)."""
    th = p[0]
    lam, mu = lame(E, nu)
    M, C, K = np.zeros((20, 20)), np.zeros((20, 20)), np.zeros((20, 20))
    cs = (lam + 2 * mu) / 4 / mu / (lam + mu)   # compliance coefficients of the stress unknowns
    cl = lam / 4 / mu / (lam + mu)
    for r, s in GP:
        N = shape(r, s)
        xg, yg = N @ X[:, 0], N @ X[:, 1]
        ax, ay, bx, by = stretch(xg, yg, rho, mu, lam, p)
        J = jac(X, r, s)
        w = th * abs(np.linalg.det(J))
        dN = np.linalg.inv(J) @ dshape(r, s)   # rows d/dx, d/dy
        NN = np.outer(N, N) * w
        Nx = np.outer(dN[0], N) * w            # [j, k] = dN_j/dx N_k
        Ny = np.outer(dN[1], N) * w
        for coef, A in ((ax * ay, M), (ax * by + ay * bx, C), (bx * by, K)):
            blk = [(0, 0, rho), (1, 1, rho), (2, 2, -cs), (3, 3, -cs), (4, 4, -1 / mu), (2, 3, cl), (3, 2, cl)]
            for i, j, v in blk:
                A[i::5, j::5] += coef * v * NN
        # Coupling of displacements and stresses: damping with ax, ay, stiffness with bx, by.
        for (ex, ey), A in (((ax, ay), C), ((bx, by), K)):
            A[0::5, 2::5] += ey * Nx
            A[2::5, 0::5] += ey * Nx.T
            A[0::5, 4::5] += ex * Ny
            A[4::5, 0::5] += ex * Ny.T
            A[1::5, 3::5] += ex * Ny
            A[3::5, 1::5] += ex * Ny.T
            A[1::5, 4::5] += ey * Nx
            A[4::5, 1::5] += ey * Nx.T
    return M, C, K


def quad_matrices(X, E, nu, rho, th, lumped):
    """K, M (8 x 8) of lin2DQuad4 in plane strain (lin2DQuad4.cpp:243-318)."""
    lam, mu = lame(E, nu)
    D = np.array([[lam + 2 * mu, lam, 0], [lam, lam + 2 * mu, 0], [0, 0, mu]])
    K, M = np.zeros((8, 8)), np.zeros((8, 8))
    for r, s in GP:
        N = shape(r, s)
        J = jac(X, r, s)
        w = th * abs(np.linalg.det(J))
        dN = np.linalg.inv(J) @ dshape(r, s)
        B = np.zeros((3, 8))
        B[0, 0::2], B[1, 1::2] = dN[0], dN[1]
        B[2, 0::2], B[2, 1::2] = dN[1], dN[0]
        K += w * B.T @ D @ B
        H = np.zeros((2, 8))
        H[0, 0::2], H[1, 1::2] = N, N
        M += w * rho * H.T @ H
    if lumped:
        M = np.diag(M.sum(axis=1))
    return K, M


@dataclass
class MixedModel:
    """Nodes with 2 or 5 unknowns; elements ('quad' | 'pml', conn, E, nu, rho, params);
    EQUAL constraints (slave node, slave unknown, master node, master unknown, factor)."""
    coords: np.ndarray
    ndof: list
    fixed: list                       # per node: list of booleans, one per unknown
    elems: list = field(default_factory=list)
    equal: list = field(default_factory=list)
    lumped: bool = False
    loads: list = field(default_factory=list)   # (node, unknown, factor, series)

    def numbering(self):
        off = np.concatenate([[0], np.cumsum(self.ndof)]).astype(int)
        slaves = {(s, a): (mn, b, f) for s, a, mn, b, f in self.equal}
        free = {}
        nfree = 0
        for n in range(len(self.ndof)):
            for a in range(self.ndof[n]):
                if (n, a) in slaves or self.fixed[n][a]:
                    continue
                free[(n, a)] = nfree
                nfree += 1
        ntot = int(off[-1])
        T = np.zeros((ntot, nfree))
        for (n, a), f in free.items():
            T[off[n] + a, f] = 1.0
        for (s, a), (mn, b, fac) in slaves.items():
            if (mn, b) in free:
                T[off[s] + a, free[(mn, b)]] = fac
        return off, T

    def assemble(self):
        off, T = self.numbering()
        ntot = T.shape[0]
        M, C, K = np.zeros((ntot, ntot)), np.zeros((ntot, ntot)), np.zeros((ntot, ntot))
        for kind, conn, E, nu, rho, p in self.elems:
            X = self.coords[list(conn)]
            if kind == "pml":
                Me, Ce, Ke = pml_matrices(X, E, nu, rho, p)
                dofs = [off[n] + a for n in conn for a in range(5)]
            else:
                Ke, Me = quad_matrices(X, E, nu, rho, p[0], self.lumped)
                Ce = np.zeros_like(Ke)
                dofs = [off[n] + a for n in conn for a in range(2)]
            ix = np.ix_(dofs, dofs)
            M[ix] += Me
            C[ix] += Ce
            K[ix] += Ke
        return off, T, M, C, K

    def keff_matrix(self, dt):
        off, T, M, C, K = self.assemble()
        return T.T @ (K + 4 / dt ** 2 * M + 2 / dt * C) @ T

    def keff_solver(self, dt):
        """Dense solve of Keff x = b after symmetric scaling by |diag|^-1/2. Keff mixes
        rho/dt^2 on the displacements with compliance/dt^2 on the stress unknowns, eleven
        orders of magnitude apart: LU of the raw matrix (condition ~5e12 on SVL's F11 column)
        loses ~1e-11 per solve, the scaled one (condition ~20) keeps ~1e-16."""
        Keff = self.keff_matrix(dt)
        S = 1.0 / np.sqrt(np.abs(np.diag(Keff)))
        As = Keff * S[:, None] * S[None, :]
        return lambda b: S * np.linalg.solve(As, S * b)

    def run_newmark(self, dt, nt, rec, k0=1, uva=False):
        """ExtendedNewmarkBeta (G = 0) + Linear, steps k = k0..nt-1 with load sample k; returns
        the total-DOF histories of `rec` [nt-k0, len(rec)] (row k-k0 after step k), or the
        (U, V, A) histories when `uva`."""
        off, T, M, C, K = self.assemble()
        ntot = T.shape[0]
        solve = self.keff_solver(dt)
        U, V, A = np.zeros(ntot), np.zeros(ntot), np.zeros(ntot)
        idx = [off[n] + a for n, a in rec]
        out = np.zeros((3, nt - k0, len(rec)))
        for k in range(k0, nt):
            F = np.zeros(ntot)
            for n, a, fac, series in self.loads:
                F[off[n] + a] += fac * (series[k] if k < len(series) else 0.0)
            Feff = T.T @ (F - K @ U + C @ V + M @ (4 / dt * V + A))
            dU = T @ solve(Feff)
            U = U + dU
            A = 4 / dt ** 2 * dU - 4 / dt * V - A
            V = 2 / dt * dU - V
            out[:, k - k0] = U[idx], V[idx], A[idx]
        return out if uva else out[0]
