"""GPU tests of independent meshes advanced together (svlg_cd_init_batch,
svlg_cd_graph_steps, svlg_autotune_tiles).

Bars: every part of the union, stepped by one fused launch per step with its own dt, matches
the oracle run of that part alone to relative L2 <= 1e-10; hipGraph replays (several, with
eager steps and synchronisations between them) give the bitwise result of eager stepping; the
autotuner returns a buildable candidate whose model reproduces the same history.
"""
import numpy as np
import pytest

import oracle
from svl_amd import batch as B
from svl_amd import workloads as W
from svl_amd.integrators import DeviceModel

pytestmark = pytest.mark.gpu

REL_TOL = 1e-10


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_batch(nt=161):
    return W.batch2d(3, nt=nt, nx_range=(8, 20), ny_range=(5, 12))


def device(u, nptr, fptr, dts, tile_nodes=0):
    dm = DeviceModel(u, tile_nodes=tile_nodes)
    cd = B.BatchedCentralDifference(dm, dts, nptr, fptr)
    cd.Initialize()
    return dm, cd


@pytest.mark.parametrize("tile_nodes", [40, 0])
def test_parts_match_independent_oracle_runs(tile_nodes):
    models, dts, nt = small_batch()
    u, nptr, fptr = B.union(models)
    recs, refs = [], []
    for i, m in enumerate(models):
        r = list(range(0, m.nfree, 7))
        refs.append(oracle.run(m, oracle.CENTRAL_DIFFERENCE, dts[i], nt, r)["U"])
        recs.append([fptr[i] + d for d in r])
    dm, cd = device(u, nptr, fptr, dts, tile_nodes)
    hist = cd.advance(1, nt - 1, rec_dofs=np.concatenate(recs))
    off = 0
    for i, ref in enumerate(refs):
        h = hist[:, off:off + ref.shape[1]]
        off += ref.shape[1]
        err = rel_l2(h, ref)
        print(f"part {i}: dt {dts[i]:.4e}, nfree {models[i].nfree}, rel L2 {err:.2e}")
        assert err < REL_TOL
    dm.close()


def test_graph_replays_equal_eager_steps():
    models, dts, nt = small_batch(nt=400)
    u, nptr, fptr = B.union(models)
    dm, cd = device(u, nptr, fptr, dts)
    cd.advance(1, 99)
    cd.advance(100, 3)
    cd.advance(103, 198)
    Ue = cd.GetDisplacements()
    dm.close()
    dm, cd = device(u, nptr, fptr, dts)
    cd.advance(1, 99, graph_steps=30)   # 3 replays + 9 eager steps
    cd.advance(100, 3)                   # eager between replays (ring slot preserved)
    cd.advance(103, 198, graph_steps=30)  # graph reused: 6 replays + 18 eager
    Ug = cd.GetDisplacements()
    assert np.array_equal(Ug, Ue)
    # And against the oracle for one part.
    m = models[1]
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dts[1], 301, [])
    assert rel_l2(cd.GetDisplacements(1), ref["final_U"]) < REL_TOL
    dm.close()


def test_autotuner_picks_a_working_tile_size():
    models, dts, nt = small_batch(nt=61)
    u, nptr, fptr = B.union(models)
    best, times = B.autotune_tile_nodes(u, nptr, dts, candidates=(32, 64, 128, 100000), steps=5)
    print("autotune:", best, times)
    assert best in (32, 64, 128, 100000)
    assert times[best] is not None and times[best] == min(t for t in times.values() if t is not None)
    dm, cd = device(u, nptr, fptr, dts, tile_nodes=best)
    cd.advance(1, 60)
    dm2, cd2 = device(u, nptr, fptr, dts, tile_nodes=32)
    cd2.advance(1, 60)
    assert rel_l2(cd.GetDisplacements(), cd2.GetDisplacements()) < 1e-12
    dm.close()
    dm2.close()
