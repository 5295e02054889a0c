"""The integrators' CG solves replayed from captured hipGraphs (runs of eight iterations while
they stay within the previous solve's count, the rest of that count as one graph, pairs, one
iteration at the predicted end) against
the same solves launched eagerly (svlg_exec_opts.cg_launch = 1): the same kernels in the same
iteration order, so the displacement histories, the iteration counts and the Newton residuals
must agree bit for bit. Jacobi-PCG solves of the 12x12x10 block take about 40 iterations (runs,
pairs and the single tail all occur), the learned flexible CG about 6.
"""
import numpy as np
import pytest

from svl_amd import workloads as W
from svl_amd.integrators import DeviceModel, ExtendedNewmarkBeta, LearnedPreconditioner, LinearSystem, NewmarkBeta

pytestmark = pytest.mark.gpu


def newmark_history(m, dt, nt, cg_launch, learned):
    dm = DeviceModel(m, cg_launch=cg_launch)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-10, maxit=3000)
    nm.Initialize()
    if learned:
        pc = LearnedPreconditioner(nm, channels=1, layers=6, mfma=-1)
        pc.use(True)
    rec = list(range(0, m.nfree, 7))
    hist = nm.advance(1, nt - 1, rec)
    its = nm.iters.copy()
    dm.close()
    return hist, its


@pytest.mark.parametrize("learned", [False, True])
def test_newmark_graph_replays_match_eager_launches(learned):
    m, dt, nt = W.block3d_implicit(12, 12, 10, nt=31)
    h_g, it_g = newmark_history(m, dt, nt, 0, learned)
    h_e, it_e = newmark_history(m, dt, nt, 1, learned)
    print(f"{'learned FCG' if learned else 'Jacobi-PCG'}: iterations per step {it_g.min()}-{it_g.max()} "
          f"(graph) vs {it_e.min()}-{it_e.max()} (eager)")
    assert np.array_equal(it_g, it_e)
    assert np.array_equal(h_g, h_e)


def test_newton_graph_replays_match_eager_launches():
    m, dt, nt = W.soil_structure_j2(16, 16, 8, block=(4, 4, 3), nt=21)
    out = []
    for cg_launch in (0, 1):
        dm = DeviceModel(m, cg_launch=cg_launch)
        nl = NewmarkBeta(dm, dt, cnvgtol=1e-6, nmax=20, cnvgtest=5, rtol=1e-10, maxit=3000)
        nl.Initialize()
        rec = list(range(0, m.nfree, 5))
        hist = nl.advance(1, nt - 1, rec)
        out.append((hist, nl.iters.copy(), nl.cg_iters.copy()))
        dm.close()
    print(f"Newton iterations {out[0][1].sum()}, CG iterations {out[0][2].sum()} (graph) vs "
          f"{out[1][2].sum()} (eager)")
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
    assert np.array_equal(out[0][0], out[1][0])


def test_cg_launch_option_checked():
    m, dt, nt = W.block3d_implicit(4, 4, 4, nt=3)
    with pytest.raises(Exception, match="cg_launch"):
        DeviceModel(m, cg_launch=2)


@pytest.mark.parametrize("channels,layers", [(1, 2), (1, 3), (2, 2), (2, 3)])
def test_learned_stopping_test_in_first_layer(channels, layers):
    """The flexible solve applies CG's stopping test in the first layer of its preconditioner
    (with two layers that layer is also the output layer, with two channels it is the general
    layer kernel). Each solve must reach its tolerance on the true residual ||b - Keff x||,
    recomputed by svlg_keff_apply; repeated solves use the previous count as the prediction;
    graph replays must equal eager launches bit for bit."""
    m, dt, nt = W.block3d_implicit(8, 8, 6, nt=9)
    rng = np.random.default_rng(7 + 10 * channels + layers)
    hists = []
    for cg_launch in (0, 1):
        dm = DeviceModel(m, cg_launch=cg_launch)
        nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-10, maxit=3000)
        nm.Initialize()
        pc = LearnedPreconditioner(nm, channels=channels, layers=layers, mfma=-1)
        pc.use(True)
        if cg_launch == 0:
            ls = LinearSystem(nm)
            for trial in range(3):
                b = rng.standard_normal(m.nfree)
                ls.SetVector(b)
                assert not ls.SolveSystem()
                x = ls.GetSolution()
                true_rel = np.linalg.norm(b - ls.Apply(x)) / np.linalg.norm(b)
                print(f"C={channels} L={layers} solve {trial}: {ls.iters} iterations, relres {ls.relres:.2e}, "
                      f"true {true_rel:.2e}")
                assert ls.iters > 0 and ls.relres <= 1e-10
                assert true_rel <= 2e-10
        rec = list(range(0, m.nfree, 5))
        hists.append((nm.advance(1, nt - 1, rec), nm.iters.copy()))
        dm.close()
    assert np.array_equal(hists[0][1], hists[1][1])
    assert np.array_equal(hists[0][0], hists[1][0])
