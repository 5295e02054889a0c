"""PML2DQuad4 / PML3DHexa8 and the mixed-unknown models around them: a numpy restatement
(test infrastructure only; no product code calls it).

SVL couples a perfectly matched layer of PML2DQuad4 elements to lin2DQuad4 elements through
EQUAL constraints between coincident nodes. A PML node carries five unknowns (ux, uy and the
stress-like components xx, yy, xy); a regular node two. This module restates, for small
dense models:

* PML2DQuad4 mass, damping and stiffness matrices (20 x 20, node-major, five unknowns per
  node) with the stretching functions of PML2DQuad4::ComputePMLStretchingFactors
  (RP/04-Elements/06-Quadrilateral/PML2DQuad4.cpp:233-466, 655-693; parameters th, m, L, R,
  x0, y0, nx, ny in the order of the constructor, :11-12), 2 x 2 Gauss points with SVL's
  abscissa; the 2D element has no history matrix (ComputePMLMatrix is empty, :467-471);
* PML3DHexa8 mass, damping, stiffness and history matrices (72 x 72, nine unknowns per node:
  ux, uy, uz, xx, yy, zz, xy, yz, xz; PML3DHexa8.cpp:214-715, stretching :952-996 with
  parameters m, L, R, x0, y0, z0, nx, ny, nz), 2 x 2 x 2 Gauss points;
* lin2DQuad4 plane-strain and lin3DHexa8 stiffness and consistent / row-sum lumped mass;
* assembly in element order, the total-to-free map with EQUAL constraints (a slave unknown
  takes the master's free number times the factor, RP/06-Mesh/Mesh.cpp:328-383) and
  ExtendedNewmarkBeta with the Linear algorithm, the history matrix G and the integrated
  displacement Ubar (ExtendedNewmarkBeta.cpp:123-154, 208-245: Keff = K + 4/dt^2 M + 2/dt C
  + dt/3 G, Feff adds -G (Ubar + dt U + dt^2/6 V), Ubar += dt U + dt/3 dU + dt^2/6 V before
  U moves; G = 0 in 2D), solved densely after a symmetric diagonal scaling (keff_solver).

Pins: tests/test_oracle_pml.py (symmetry, closed forms without stretching, the Laplace-domain
identity of a uniform stretch in 2D and 3D, the lin2DQuad4 / lin3DHexa8 matrices against the
C++ oracle, a soil column whose PML-truncated response follows a much longer column while a
fixed truncation does not, the reflection coefficient 2R of the 2D column and the 3D rod that
fixes the damping profile b0, the accuracy of the scaled dense solve, and the equation of
motion with G Ubar plus the Hermite integral of Ubar at every step of a 3D model with G != 0).

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


GP3 = [(r, s, t) for t in (-G, G) for s in (-G, G) for r in (-G, G)]  # r fastest
CORNERS3 = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]


def shape3(r, s, t):
    return np.array([(1 + a * r) * (1 + b * s) * (1 + c * t) for a, b, c in CORNERS3]) / 8.0


def dshape3(r, s, t):
    """[dN/dr, dN/ds, dN/dt] rows of the eight corners."""
    return np.array([[a * (1 + b * s) * (1 + c * t) for a, b, c in CORNERS3],
                     [b * (1 + a * r) * (1 + c * t) for a, b, c in CORNERS3],
                     [c * (1 + a * r) * (1 + b * s) for a, b, c in CORNERS3]]) / 8.0


def stretch3(x, y, z, rho, mu, lam, p):
    """(ax, ay, az, bx, by, bz); p = (m, L, R, x0, y0, z0, nx, ny, nz) (PML3DHexa8.cpp:952-996)."""
    m, L, R, x0, y0, z0, nx, ny, nz = p
    vp = math.sqrt((lam + 2 * mu) / rho)
    a0 = (m + 1) * (L / 10.0) / 2 / L * math.log(1 / R)
    b0 = (m + 1) * vp / 2 / L * math.log(1 / R)
    f = [((x - x0) * nx / L) ** m, ((y - y0) * ny / L) ** m, ((z - z0) * nz / L) ** m]
    return [1 + a0 * v for v in f] + [b0 * v for v in f]


# Displacement-stress coupling of PML3DHexa8: (displacement, stress unknown) pairs whose entry
# is dN_j/dx_d N_k, per derivative direction d.
COUPLE3 = {0: ((0, 3), (1, 6), (2, 8)), 1: ((0, 6), (1, 4), (2, 7)), 2: ((0, 8), (1, 7), (2, 5))}


def pml3_matrices(X, E, nu, rho, p):
    """M, C, K, G (72 x 72) of PML3DHexa8 (PML3DHexa8.cpp:214-715)."""
    lam, mu = lame(E, nu)
    mats = [np.zeros((72, 72)) for _ in range(4)]
    cs = (lam + mu) / mu / (3 * lam + 2 * mu)
    cl = lam / 2 / mu / (3 * lam + 2 * mu)
    blk = [(i, i, rho) for i in range(3)] + [(i, i, -cs) for i in (3, 4, 5)] + [(i, i, -1 / mu) for i in (6, 7, 8)]
    blk += [(i, j, cl) for i in (3, 4, 5) for j in (3, 4, 5) if i != j]
    for r, s, t in GP3:
        N = shape3(r, s, t)
        xg, yg, zg = N @ X
        ax, ay, az, bx, by, bz = stretch3(xg, yg, zg, rho, mu, lam, p)
        J = dshape3(r, s, t) @ X
        w = abs(np.linalg.det(J))
        dN = np.linalg.inv(J) @ dshape3(r, s, t)
        NN = np.outer(N, N) * w
        mass_like = (ax * ay * az, ax * ay * bz + ax * by * az + bx * ay * az,
                     ax * by * bz + bx * by * az + bx * ay * bz, bx * by * bz)
        couple = (None, (ay * az, ax * az, ax * ay), (ay * bz + az * by, ax * bz + az * bx, ax * by + ay * bx),
                  (by * bz, bx * bz, bx * by))
        for A, coef, cc in zip(mats, mass_like, couple):
            for i, j, v in blk:
                A[i::9, j::9] += coef * v * NN
            if cc is None:
                continue
            for d in range(3):
                D = np.outer(dN[d], N) * w * cc[d]
                for pu, ps in COUPLE3[d]:
                    A[pu::9, ps::9] += D
                    A[ps::9, pu::9] += D.T
    return tuple(mats)


def hexa_matrices(X, E, nu, rho, lumped):
    """K, M (24 x 24) of lin3DHexa8 (lin3DHexa8.cpp)."""
    lam, mu = lame(E, nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] += 2 * mu
    D[[3, 4, 5], [3, 4, 5]] = mu
    K, M = np.zeros((24, 24)), np.zeros((24, 24))
    for r, s, t in GP3:
        N = shape3(r, s, t)
        J = dshape3(r, s, t) @ X
        w = abs(np.linalg.det(J))
        dN = np.linalg.inv(J) @ dshape3(r, s, t)
        B = np.zeros((6, 24))
        B[0, 0::3], B[1, 1::3], B[2, 2::3] = dN[0], dN[1], dN[2]
        B[3, 0::3], B[3, 1::3] = dN[1], dN[0]
        B[4, 1::3], B[4, 2::3] = dN[2], dN[1]
        B[5, 0::3], B[5, 2::3] = dN[2], dN[0]
        K += w * B.T @ D @ B
        H = np.zeros((3, 24))
        H[0, 0::3], H[1, 1::3], H[2, 2::3] = N, N, N
        M += w * rho * H.T @ H
    if lumped:
        M = np.diag(M.sum(axis=1))
    return K, M


@dataclass
class MixedModel:
    """Nodes with 2 or 5 (2D) or 3 or 9 (3D) unknowns; elements ('quad' | 'pml' | 'hexa' |
    'pml3', conn, E, nu, rho, params: (th,) / (th, m, L, R, x0, y0, nx, ny) / () /
    (m, L, R, x0, y0, z0, nx, ny, nz)); EQUAL constraints (slave node, slave unknown, master
    node, master unknown, factor)."""
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
        M, C, K, Gm = (np.zeros((ntot, ntot)) for _ in range(4))
        for kind, conn, E, nu, rho, p in self.elems:
            X = self.coords[list(conn)]
            Ge = None
            if kind == "pml":
                Me, Ce, Ke = pml_matrices(X, E, nu, rho, p)
                nun = 5
            elif kind == "pml3":
                Me, Ce, Ke, Ge = pml3_matrices(X, E, nu, rho, p)
                nun = 9
            elif kind == "hexa":
                Ke, Me = hexa_matrices(X, E, nu, rho, self.lumped)
                Ce = np.zeros_like(Ke)
                nun = 3
            else:
                Ke, Me = quad_matrices(X, E, nu, rho, p[0], self.lumped)
                Ce = np.zeros_like(Ke)
                nun = 2
            dofs = [off[n] + a for n in conn for a in range(nun)]
            ix = np.ix_(dofs, dofs)
            M[ix] += Me
            C[ix] += Ce
            K[ix] += Ke
            if Ge is not None:
                Gm[ix] += Ge
        self._G = Gm
        return off, T, M, C, K

    def keff_matrix(self, dt):
        off, T, M, C, K = self.assemble()
        return T.T @ (K + 4 / dt ** 2 * M + 2 / dt * C + dt / 3 * self._G) @ T

    def keff_solver(self, dt):
        """Dense solve of Keff x = b after symmetric scaling by |diag|^-1/2. Keff mixes
        rho/dt^2 on the displacements with compliance/dt^2 on the stress unknowns, eleven
        orders of magnitude apart: LU of the raw matrix (condition ~5e12 on SVL's F11 column)
        loses ~1e-11 per solve, the scaled one (condition ~20) keeps ~1e-16."""
        from scipy.linalg import lu_factor, lu_solve
        Keff = self.keff_matrix(dt)
        S = 1.0 / np.sqrt(np.abs(np.diag(Keff)))
        lu = lu_factor(Keff * S[:, None] * S[None, :])
        return lambda b: S * lu_solve(lu, S * b)

    def run_newmark(self, dt, nt, rec, k0=1, uva=False, full=False):
        """ExtendedNewmarkBeta + Linear, steps k = k0..nt-1 with load sample k; returns the
        total-DOF histories of `rec` [nt-k0, len(rec)] (row k-k0 after step k), or the (U, V, A)
        histories when `uva`, or every total DOF's (U, V, A, Ubar, F) when `full`."""
        off, T, M, C, K = self.assemble()
        Gm = self._G
        ntot = T.shape[0]
        solve = self.keff_solver(dt)
        U, V, A, Ub = np.zeros(ntot), np.zeros(ntot), np.zeros(ntot), np.zeros(ntot)
        idx = list(range(ntot)) if full else [off[n] + a for n, a in rec]
        out = np.zeros((5 if full else 3, nt - k0, len(idx)))
        for k in range(k0, nt):
            F = np.zeros(ntot)
            for n, a, fac, series in self.loads:
                F[off[n] + a] += fac * (series[k] if k < len(series) else 0.0)
            Feff = T.T @ (F - K @ U - Gm @ (Ub + dt * U + dt * dt / 6 * V) + C @ V + M @ (4 / dt * V + A))
            dU = T @ solve(Feff)
            Ub = Ub + dt * U + dt / 3 * dU + dt * dt / 6 * V
            U = U + dU
            A = 4 / dt ** 2 * dU - 4 / dt * V - A
            V = 2 / dt * dU - V
            if full:
                out[:, k - k0] = U, V, A, Ub, F
            else:
                out[:, k - k0] = U[idx], V[idx], A[idx]
        return out if (uva or full) else out[0]
