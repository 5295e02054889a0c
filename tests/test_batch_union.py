"""Host-side checks of the batched workload (no GPU): the union model keeps every part's
Plain numbering as a contiguous, offset range, its loads, dashpots and materials, and the
realization generator follows its seeds."""
import numpy as np

from svl_amd import batch as B
from svl_amd import workloads as W


def small_batch(nt=40):
    return W.batch2d(3, nt=nt, nx_range=(5, 9), ny_range=(4, 7))


def test_union_numbering_and_arrays():
    models, dts, nt = small_batch()
    u, nptr, fptr = B.union(models)
    fd = u.free_dof()
    assert u.nfree == sum(m.nfree for m in models) == fptr[-1]
    for i, m in enumerate(models):
        part = fd[nptr[i]:nptr[i + 1]]
        assert np.array_equal(part[part >= 0], m.free_dof()[m.free_dof() >= 0] + fptr[i])
        e0 = sum(q.nelem for q in models[:i])
        assert np.array_equal(u.conn[e0:e0 + m.nelem], m.conn + nptr[i])
        assert np.array_equal(u.E[e0:e0 + m.nelem], m.E)
    assert u.load_node.size == sum(m.load_node.size for m in models)
    assert u.dash_node.size == sum(m.dash_node.size for m in models)
    # Bounding boxes of the parts are disjoint along x.
    xs = [(m.coords[:, 0].min(), m.coords[:, 0].max()) for m in models]
    assert all(xs[i][1] < xs[i + 1][0] for i in range(len(xs) - 1))


def test_realizations_follow_their_seeds():
    a, da, _ = W.batch2d(4, nt=10, nx_range=(5, 40), ny_range=(5, 20))
    b, db, _ = W.batch2d(4, nt=10, nx_range=(5, 40), ny_range=(5, 20))
    assert da == db
    for x, y in zip(a, b):
        assert np.array_equal(x.E, y.E) and np.array_equal(x.load_series, y.load_series)
    sizes = {(m.conn.shape[0]) for m in a}
    assert len(sizes) > 1  # uneven sizes
    # Base forces: 2 eta_s of the bottom dashpot in x, one per bottom node.
    m = a[0]
    assert np.all(m.load_dir[:, 0] == 1.0) and np.all(m.load_factor > 0)


def test_union_carries_domain_reduction_loads():
    """Two DRM-loaded Quad4 models in one union: the oracle's run of the union equals the runs
    of the parts (one DRM load over both layers, indices offset per part)."""
    import oracle
    from svl_amd import drm
    from _meshes import drm_quad_model
    dt, nt = 1e-3, 60
    parts = []
    for scale in (1.0, 0.8):
        m, box = drm_quad_model(nt, dt)
        m.E = m.E * scale
        bg = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, list(range(m.nfree)))
        m.load_node = m.load_node[:0]
        m.load_dir, m.load_factor, m.load_series = np.zeros((0, 2)), np.zeros(0), np.zeros((0, 1))
        e, n, c = drm.drm_box(m, *box)
        drm.attach(m, e, n, c, drm.signals_from_history(m, n, nt - 10 * len(parts), bg["U"], lag=2), 1.5)
        parts.append(m)
    u, nptr, fptr = B.union(parts)
    assert u.drm_elem.size == sum(p.drm_elem.size for p in parts)
    assert u.drm_signal.shape[1] == nt
    whole = oracle.run(u, oracle.CENTRAL_DIFFERENCE, dt, nt, list(range(u.nfree)))["U"]
    for i, p in enumerate(parts):
        own = oracle.run(p, oracle.CENTRAL_DIFFERENCE, dt, nt, list(range(p.nfree)))["U"]
        assert np.abs(own).max() > 0
        assert np.array_equal(whole[:, fptr[i]:fptr[i + 1]], own)
