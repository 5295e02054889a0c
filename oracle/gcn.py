"""FP64 numpy restatement of the graph-convolution preconditioner (test infrastructure).

The network (no SVL counterpart; it stands in for the factorisation of the effective
stiffness in LinearSystem::SolveSystem, RP/11-Solvers/LinearSystem.hpp:44-66):

    x0 = s * r,  s = D^-1/2 (D = diag Keff),  Ehat = diag(s) Keff diag(s)
    H^1 = x0 c_1^T
    H^l = H^(l-1) A_l + (Ehat H^(l-1)) B_l + x0 c_l^T     l = 2..L
    z = s * (H^L w)

theta = for l = 1..L [A_l (C x C, row = input channel) | B_l (C x C) | c_l (C)], then w (C).
Loss of a pair (r, e) with Keff e = r: (z - e)^T Keff (z - e) / (e^T r); a batch takes the mean.
The backward pass here is the analytic one; the tests pin it by finite differences and
then hold the device kernels to it.
"""
from __future__ import annotations

import numpy as np


def layout(C: int, L: int):
    TL = 2 * C * C + C
    return TL, L * TL + C


def unpack(theta, C: int, L: int):
    TL, n = layout(C, L)
    theta = np.asarray(theta, dtype=np.float64)
    assert theta.size == n
    A, B, c = [], [], []
    for l in range(L):
        blk = theta[l * TL:(l + 1) * TL]
        A.append(blk[:C * C].reshape(C, C))
        B.append(blk[C * C:2 * C * C].reshape(C, C))
        c.append(blk[2 * C * C:])
    w = theta[L * TL:]
    return A, B, c, w


def scaling(K: np.ndarray) -> np.ndarray:
    return 1.0 / np.sqrt(np.diag(K))


def forward(theta, C: int, L: int, K: np.ndarray, r: np.ndarray):
    """Returns z and the intermediate states (x0, H list, Y list) for one residual r."""
    A, B, c, w = unpack(theta, C, L)
    s = scaling(K)
    Eh = (s[:, None] * K) * s[None, :]
    x0 = s * r
    H = [None, np.outer(x0, c[0])]
    Y = [None, None]
    for l in range(2, L + 1):
        y = Eh @ H[l - 1]
        Y.append(y)
        H.append(H[l - 1] @ A[l - 1] + y @ B[l - 1] + np.outer(x0, c[l - 1]))
    z = s * (H[L] @ w)
    return z, (x0, H, Y, Eh, s)


def apply(theta, C: int, L: int, K: np.ndarray, r: np.ndarray) -> np.ndarray:
    return forward(theta, C, L, K, r)[0]


def loss_grad(theta, C: int, L: int, K: np.ndarray, R: np.ndarray, E: np.ndarray):
    """Mean loss over the pairs (rows of R, E) and its gradient w.r.t. theta."""
    A, B, c, w = unpack(theta, C, L)
    TL, n = layout(C, L)
    grad = np.zeros(n)
    total = 0.0
    npairs = R.shape[0]
    for r, e in zip(R, E):
        z, (x0, H, Y, Eh, s) = forward(theta, C, L, K, r)
        kz = K @ z
        den = e @ r
        total += (z - e) @ (kz - r) / den / npairs
        gz = 2.0 * (kz - r) / den / npairs
        sg = s * gz
        grad[L * TL:] += H[L].T @ sg
        G = np.outer(sg, w)
        for l in range(L, 1, -1):
            off = (l - 1) * TL
            grad[off:off + C * C] += (H[l - 1].T @ G).reshape(-1)
            grad[off + C * C:off + 2 * C * C] += (Y[l].T @ G).reshape(-1)
            grad[off + 2 * C * C:off + TL] += x0 @ G
            G = G @ A[l - 1].T + Eh @ (G @ B[l - 1].T)
        grad[2 * C * C:TL] += x0 @ G
    return total, grad


def chebyshev_theta(L: int, lmin: float, lmax: float) -> np.ndarray:
    """One-channel parameters of L Richardson steps with reciprocal Chebyshev-node steps
    (the order does not change the product in exact arithmetic)."""
    TL, n = layout(1, L)
    th = np.zeros(n)
    for l in range(L):
        node = 0.5 * (lmax + lmin) + 0.5 * (lmax - lmin) * np.cos(np.pi * (2 * l + 1) / (2 * L))
        om = 1.0 / node
        th[l * TL + 0] = 1.0
        th[l * TL + 1] = -om
        th[l * TL + 2] = om
    th[L * TL] = 1.0
    return th
