"""GPU tests of the learned graph-convolution preconditioner.

Bars: the FP64-element forward pass against the numpy restatement (oracle/gcn.py) to 1e-5
relative (hidden states are FP32 on the device); the BF16 MFMA element step against the FP64
one to 3e-2; loss to 1e-5 and gradient to 1e-4 relative against the restatement's analytic
gradient (itself pinned by finite differences on the CPU). Training must lower the loss,
flexible CG with the network must need at most a third of the iterations of Jacobi-PCG, training
must never leave it worse than its (Chebyshev) starting point, and the Newmark history it
produces must match the oracle's direct solves to 1e-10. A strict gain of training over the
Chebyshev start is expected to fail (xfail, with the reason).
"""
import numpy as np
import pytest

import oracle
from oracle import gcn
from svl_amd import workloads as W
from svl_amd.integrators import DeviceModel, ExtendedNewmarkBeta, LearnedPreconditioner

from _meshes import block_model

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def setup(m, dt, C, L, mfma=0, theta=None, tile_nodes=0):
    dm = DeviceModel(m, tile_nodes=tile_nodes)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-10)
    nm.Initialize()
    pc = LearnedPreconditioner(nm, channels=C, layers=L, mfma=mfma, theta=theta)
    return dm, nm, pc


def with_rayleigh(m, am=0.5712, ak=1.447e-3):
    m.ray_am = np.full(m.nelem, am)
    m.ray_ak = np.full(m.nelem, ak)
    return m


CASES = {
    "block": lambda: W.block3d_implicit(6, 6, 5, nt=11)[:2],
    "distorted": lambda: (with_rayleigh(block_model(5, 4, 4, distort=0.15, seed=3)[0]), 1e-3),
    "quad": lambda: (with_rayleigh(W.layer2d(16, 8, nt=11)[0], 0.3, 2e-3), 2e-3),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("C,L", [(1, 3), (2, 4)])
def test_forward_matches_restatement(case, C, L):
    m, dt = CASES[case]()
    K = oracle.keff_dense(m, dt)
    theta = np.random.default_rng(C + L).standard_normal(gcn.layout(C, L)[1]) * 0.3
    dm, nm, pc = setup(m, dt, C, L, mfma=0, theta=theta, tile_nodes=60)
    r = np.random.default_rng(9).standard_normal(m.nfree)
    assert rel(pc.apply(r), gcn.apply(theta, C, L, K, r)) < 1e-5
    dm.close()


@pytest.mark.parametrize("C,L", [(1, 5), (2, 4)])
def test_mfma_element_step(C, L):
    """BF16 MFMA layers (the one-channel inference kernel, and the general kernel) against the
    FP64 element routine."""
    m, dt = CASES["block"]()
    dm, nm, pc = setup(m, dt, C, L, mfma=1)
    assert pc.info()["mfma"] == 1
    r = np.random.default_rng(3).standard_normal(m.nfree)
    z64 = pc.apply(r, mfma=False)
    z16 = pc.apply(r, mfma=True)
    assert rel(z16, z64) < 3e-2
    dm.close()


@pytest.mark.parametrize("case", ["block", "distorted", "quad"])
def test_loss_and_gradient(case):
    m, dt = CASES[case]()
    K = oracle.keff_dense(m, dt)
    C, L = 2, 3
    theta = np.random.default_rng(21).standard_normal(gcn.layout(C, L)[1]) * 0.3
    dm, nm, pc = setup(m, dt, C, L, mfma=0, theta=theta, tile_nodes=50)
    rng = np.random.default_rng(5)
    E = rng.standard_normal((3, m.nfree))
    R = E @ K.T
    loss, g = pc.loss_grad(R, E)
    loss_ref, g_ref = gcn.loss_grad(theta, C, L, K, R, E)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert rel(g, g_ref) < 1e-4
    dm.close()


def test_chebyshev_initialisation():
    m, dt = CASES["block"]()
    K = oracle.keff_dense(m, dt)
    s = 1.0 / np.sqrt(np.diag(K))
    lam = np.linalg.eigvalsh((s[:, None] * K) * s[None, :])
    dm, nm, pc = setup(m, dt, 1, 5, mfma=0)
    info = pc.info()
    # Lanczos Ritz values lie inside the spectrum and are close to its ends.
    assert lam[-1] * 0.99 <= info["lmax"] <= lam[-1] * (1 + 1e-9)
    assert lam[0] * (1 - 1e-9) <= info["lmin"] <= lam[0] * 1.5
    # Same polynomial as the restatement's Chebyshev parameters on the device's interval.
    r = np.random.default_rng(2).standard_normal(m.nfree)
    th = gcn.chebyshev_theta(5, 0.9 * info["lmin"], 1.05 * info["lmax"])
    assert rel(pc.apply(r), gcn.apply(th, 1, 5, K, r)) < 1e-5
    dm.close()


def newmark_iters(m, dt, nsteps, learned=None, rtol=1e-8):
    dm = DeviceModel(m)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=rtol, maxit=2000)
    nm.Initialize()
    out = None
    if learned is not None:
        pc = LearnedPreconditioner(nm, **learned["init"])
        if learned.get("train"):
            pc.train(**learned["train"])
        pc.use(True)
        out = dict(pc.info(), loss_hist=pc.loss_hist, accepted=getattr(pc, "accepted", None))
    nm.advance(1, nsteps)
    it = nm.iters.copy()
    U = nm.GetDisplacements()
    dm.close()
    return it, U, out


def trained_vs_start(m, dt):
    init = {"channels": 1, "layers": 6, "mfma": -1}
    it_j, U_j, _ = newmark_iters(m, dt, 30)
    it_c, U_c, _ = newmark_iters(m, dt, 30, {"init": init})
    it_l, U_l, pc = newmark_iters(m, dt, 30, {"init": init, "train": {"nsteps": 24, "every": 3, "sweeps": 2,
                                                                      "epochs": 60, "lr": 2e-3}})
    print(f"FCG iterations per step: Jacobi {it_j.mean():.2f}, Chebyshev-initialised network "
          f"{it_c.mean():.2f}, trained {it_l.mean():.2f} (accepted {pc['accepted']}, validation "
          f"{pc['val_iters_before']} -> {pc['val_iters_after']}, loss {pc['loss_hist'][0]:.4f} -> "
          f"{pc['loss_hist'][-1]:.4f})")
    return it_j, U_j, it_c, U_c, it_l, U_l, pc


def test_network_gain_over_jacobi_pcg():
    """What the network gains over the method without it (Jacobi-PCG, the fallback and judge):
    at most a third of the iterations per step, the same solution to the solver tolerance, and
    training lowers its loss and never leaves the solver worse than its start: the validation
    guard keeps trained parameters only when they cut the validation count, and otherwise
    restores the start, whose iteration counts then repeat exactly."""
    m, dt, nt = W.block3d_implicit(12, 12, 10, nt=61)
    it_j, U_j, it_c, U_c, it_l, U_l, pc = trained_vs_start(m, dt)
    assert it_l.mean() <= it_j.mean() / 3.0 and it_c.mean() <= it_j.mean() / 3.0
    assert pc["loss_hist"][-1] < pc["loss_hist"][0]
    if pc["accepted"]:
        assert pc["val_iters_after"] < pc["val_iters_before"]
    else:
        assert np.array_equal(it_l, it_c)
    assert it_l.mean() <= it_c.mean()
    assert rel(U_l, U_j) < 1e-6 and rel(U_c, U_j) < 1e-6


@pytest.mark.xfail(reason="the network is a polynomial in Ehat of fixed degree, and its Chebyshev start is "
                          "already near the best such polynomial for this spectrum (least-squares polynomials "
                          "of the same degree need as many or more FCG iterations), so training keeps the start",
                   strict=False)
def test_training_gain_over_chebyshev_start():
    m, dt, nt = W.block3d_implicit(12, 12, 10, nt=61)
    it_j, U_j, it_c, U_c, it_l, U_l, pc = trained_vs_start(m, dt)
    assert pc["accepted"]
    assert it_l.mean() < it_c.mean()


def test_learned_fcg_history_matches_oracle():
    m, dt, nt = W.block3d_implicit(8, 8, 6, nt=41)
    rec = W.dofs_of(m, int(m.load_node[0])) + list(range(0, m.nfree, 19))
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec)
    dm = DeviceModel(m)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=3000)
    nm.Initialize()
    pc = LearnedPreconditioner(nm, channels=2, layers=5, mfma=-1)
    pc.train(nsteps=12, every=3, sweeps=1, epochs=20, lr=2e-3)
    pc.use(True)
    hist = nm.advance(1, nt - 1, rec)
    assert np.all(nm.relres <= 1e-13)
    assert rel(hist, ref["U"]) < 1e-10
    dm.close()
