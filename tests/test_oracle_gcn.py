"""Pins the numpy restatement of the graph-convolution preconditioner (oracle/gcn.py).

* its analytic gradient against central finite differences of its own loss;
* special cases with closed forms: the Jacobi preconditioner (z = D^-1 r) and the
  Chebyshev initialisation, whose residual polynomial 1 - lambda q(lambda) must equal the
  scaled Chebyshev polynomial T_L((theta - lambda)/delta) / T_L(theta/delta) on every
  eigenvalue of Ehat;
* the loss equals the energy-norm error ratio of an independent evaluation.
"""
import numpy as np
import pytest

import oracle
from oracle import gcn
from svl_amd import workloads as W


@pytest.fixture(scope="module")
def small():
    m, dt, nt = W.block3d_implicit(3, 3, 2, nt=11)
    K = oracle.keff_dense(m, dt)
    return m, dt, K


def pairs(K, n, seed):
    rng = np.random.default_rng(seed)
    E = rng.standard_normal((n, K.shape[0]))
    R = E @ K.T
    return R, E


def test_keff_dense_matches_operator(small):
    m, dt, K = small
    x = np.random.default_rng(1).standard_normal(m.nfree)
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())
    y = oracle.keff_apply(m, dt, x)
    assert np.linalg.norm(K @ x - y) <= 1e-14 * np.linalg.norm(y)


@pytest.mark.parametrize("C,L", [(1, 2), (2, 3), (3, 4)])
def test_gradient_finite_differences(small, C, L):
    m, dt, K = small
    TL, n = gcn.layout(C, L)
    rng = np.random.default_rng(C * 10 + L)
    theta = rng.standard_normal(n) * 0.3
    R, E = pairs(K, 2, 5)
    loss, g = gcn.loss_grad(theta, C, L, K, R, E)
    fd = np.zeros(n)
    h = 1e-6
    for i in range(n):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (gcn.loss_grad(tp, C, L, K, R, E)[0] - gcn.loss_grad(tm, C, L, K, R, E)[0]) / (2 * h)
    assert np.linalg.norm(g - fd) <= 1e-6 * np.linalg.norm(fd)


def test_jacobi_special_case(small):
    m, dt, K = small
    L = 3
    TL, n = gcn.layout(1, L)
    theta = np.zeros(n)
    for l in range(L):
        theta[l * TL] = 1.0          # A_l = 1, B_l = 0
    theta[2] = 1.0                   # c_1 = 1
    theta[L * TL] = 1.0              # w = 1
    r = np.random.default_rng(2).standard_normal(m.nfree)
    z = gcn.apply(theta, 1, L, K, r)
    assert np.allclose(z, r / np.diag(K), rtol=1e-14, atol=0)


def test_chebyshev_residual_polynomial(small):
    m, dt, K = small
    s = 1.0 / np.sqrt(np.diag(K))
    Eh = (s[:, None] * K) * s[None, :]
    lam, V = np.linalg.eigh(Eh)
    lmin, lmax = lam[0], lam[-1]
    for L in (2, 4, 7):
        theta = gcn.chebyshev_theta(L, lmin, lmax)
        # Network response to each eigenvector v (r = v / s so that x0 = v): z = s q(lambda) v.
        q = np.array([(gcn.apply(theta, 1, L, K, V[:, i] / s) / s) @ V[:, i] for i in range(0, lam.size, 7)])
        lsub = lam[::7]
        th, de = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
        cheb = np.polynomial.chebyshev.Chebyshev.basis(L)
        expect = cheb((th - lsub) / de) / cheb(th / de)
        assert np.allclose(1.0 - lsub * q, expect, atol=1e-12)


def test_loss_is_energy_error_ratio(small):
    m, dt, K = small
    C, L = 2, 3
    theta = np.random.default_rng(4).standard_normal(gcn.layout(C, L)[1]) * 0.2
    R, E = pairs(K, 1, 8)
    z = gcn.apply(theta, C, L, K, R[0])
    d = z - E[0]
    Lc = np.linalg.cholesky(K)
    expect = np.sum((Lc.T @ d) ** 2) / np.sum((Lc.T @ E[0]) ** 2)
    assert abs(gcn.loss_grad(theta, C, L, K, R, E)[0] - expect) <= 1e-12 * expect
