"""GPU parity of the domain reduction load (ElementLoad GENERALWAVE).

Bars:
  * displacement histories of a DRM-loaded model against the oracle's run of the same model
    (lin*::ComputeDomainReductionForces restated with dense masked element matrices):
    relative L2 <= 1e-10 for CentralDifference (fused corner-slot and element-colored step
    kernels) and for ExtendedNewmarkBeta with the CG tolerance at 1e-13;
  * reactions at fixed nodes of the DRM layer (whose Fext holds DRM forces) against the
    oracle's, relative to the largest reaction: <= 1e-9 (signals scaled per node, so that
    the reactions do not cancel to round-off);
  * graph-captured CentralDifference steps give the eager steps' state bit for bit;
  * at a size the oracle does not run (a 40x40x20 Hexa8 block): the device's own free field
    from a background run, fed back as DRM signals, is reproduced inside the box and the
    exterior stays at rest, both to 1e-9 of the largest displacement.
"""
import numpy as np
import pytest

import oracle
from svl_amd import _capi as capi, drm
from svl_amd.integrators import CentralDifference, DeviceModel, ExtendedNewmarkBeta
from svl_amd.mesh import Model

from _meshes import drm_hexa_model, drm_quad_model

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def reduced(m, box, nt, U, V=None, A=None, lag=1, cols=None):
    elems, nodes, cond = drm.drm_box(m, *box)
    r = Model(**{k: getattr(m, k) for k in ("ndim", "coords", "node_tags", "elem_type", "conn", "E", "nu", "rho",
                                             "restrained", "ray_am", "ray_ak", "massform_lumped")})
    drm.attach(r, elems, nodes, cond, drm.signals_from_history(m, nodes, nt, U, V, A, lag, cols))
    return r


def inside_dofs(m, box):
    fd = m.free_dof()
    x0, xl = (np.asarray(v, float) for v in box)
    inside = np.all(np.abs(m.coords - x0) <= xl, axis=1)
    ins = np.array([fd[n, c] for n in np.nonzero(inside)[0] for c in range(m.ndim) if fd[n, c] >= 0])
    out = np.array([fd[n, c] for n in np.nonzero(~inside)[0] for c in range(m.ndim) if fd[n, c] >= 0])
    return ins, out


def case(kind, integ):
    if integ == oracle.CENTRAL_DIFFERENCE:
        dt, nt, lag, am, ak = 1e-3, 200, 2, 0.3, 0.0
    else:
        dt, nt, lag, am, ak = 2e-3, 100, 1, 0.4, 3e-4
    m, box = (drm_quad_model if kind == "quad" else drm_hexa_model)(nt, dt)
    m.ray_am = np.full(m.nelem, am)
    m.ray_ak = np.full(m.nelem, ak)
    rec = list(range(m.nfree))
    bg = oracle.run(m, integ, dt, nt, rec, want_va=True)
    r = reduced(m, box, nt, bg["U"], bg["V"], bg["A"], lag)
    return m, r, box, dt, nt, rec, bg


@pytest.mark.parametrize("variant", [3, 2])
@pytest.mark.parametrize("kind", ["quad", "hexa"])
def test_central_difference_matches_oracle(kind, variant):
    m, r, box, dt, nt, rec, bg = case(kind, oracle.CENTRAL_DIFFERENCE)
    ref = oracle.run(r, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    dm = DeviceModel(r, variant=variant)
    cd = CentralDifference(dm, dt)
    cd.Initialize()
    hist = cd.advance(1, nt - 1, rec)
    assert rel_l2(hist, ref["U"]) <= 1e-10
    ins, out = inside_dofs(m, box)
    scale = np.abs(bg["U"]).max()
    assert np.abs(hist[:, ins] - bg["U"][:, ins]).max() <= 1e-9 * scale
    assert np.abs(hist[:, out]).max() <= 1e-9 * scale


@pytest.mark.parametrize("kind", ["quad", "hexa"])
def test_newmark_matches_oracle(kind):
    m, r, box, dt, nt, rec, bg = case(kind, oracle.EXTENDED_NEWMARK)
    ref = oracle.run(r, oracle.EXTENDED_NEWMARK, dt, nt, rec, want_va=True)
    dm = DeviceModel(r)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=5000)
    nm.Initialize()
    hist = nm.advance(1, nt - 1, rec)
    assert rel_l2(hist, ref["U"]) <= 1e-10
    assert rel_l2(nm.GetVelocities(), ref["V"][-1]) <= 1e-9
    ins, out = inside_dofs(m, box)
    scale = np.abs(bg["U"]).max()
    assert np.abs(hist[:, ins] - bg["U"][:, ins]).max() <= 1e-9 * scale


def test_reactions_on_drm_layer_match_oracle():
    # A box reaching x = 0.5 .. 15.5: the exterior DRM nodes on x = 0 and x = 16 are fixed.
    # With a free field (or any scaled or shifted copy) their reactions cancel to round-off
    # (Fint = the DRM force); per-node scaled signals are no free field, which leaves
    # reactions of the size of the DRM forces.
    dt, nt = 2e-3, 60
    m, box = drm_quad_model(nt, dt, half=(7.5, 2.5))
    m.ray_am = np.full(m.nelem, 0.4)
    m.ray_ak = np.full(m.nelem, 3e-4)
    rec = list(range(m.nfree))
    bg = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec, want_va=True)
    r = reduced(m, box, nt, bg["U"], bg["V"], bg["A"], 1)
    r.drm_signal = r.drm_signal * np.random.default_rng(5).uniform(0.5, 1.5, (r.drm_node.size, 1, 1))
    fixed = np.intersect1d(np.nonzero(r.restrained.any(1))[0], r.drm_node[r.drm_cond == 1])
    assert fixed.size >= 4
    tot = [int(n) * 2 + c for n in fixed for c in range(2)]
    ref = oracle.run(r, oracle.EXTENDED_NEWMARK, dt, nt, rec, reac_tot=tot)
    dm = DeviceModel(r)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=5000)
    nm.Initialize()
    nm.advance(1, nt - 1)
    R = dm.reactions(fixed, nt - 1).reshape(-1)
    want = ref["R"][nt - 2]
    assert np.abs(want).max() > 0
    assert np.abs(R - want).max() <= 1e-9 * np.abs(want).max()


def test_graph_steps_match_eager():
    m, r, box, dt, nt, rec, bg = case("quad", oracle.CENTRAL_DIFFERENCE)
    states = []
    for graph in (False, True):
        dm = DeviceModel(r)
        cd = CentralDifference(dm, dt)
        cd.Initialize()
        if graph:
            capi.check(capi.lib().svlg_cd_graph_steps(dm.h, 1, 12 * 6, 6), "svlg_cd_graph_steps")
        else:
            cd.advance(1, 12 * 6)
        states.append(cd.GetDisplacements())
    assert np.abs(states[0]).max() > 0
    assert np.array_equal(states[0], states[1])


def test_free_field_reproduced_at_size():
    dt, nt = 1e-3, 200
    m, box = drm_hexa_model(nt, dt, n=(40, 40, 20))
    elems, nodes, cond = drm.drm_box(m, *box)
    ins, out = inside_dofs(m, box)
    fd = m.free_dof()
    drm_dofs = fd[nodes].reshape(-1)
    drm_dofs = drm_dofs[drm_dofs >= 0]
    rng = np.random.default_rng(0)
    check = np.unique(np.concatenate([rng.choice(ins, 1500, replace=False), rng.choice(out, 1500, replace=False)]))
    cols = np.unique(np.concatenate([drm_dofs, check])).astype(np.int32)
    dm = DeviceModel(m)
    cd = CentralDifference(dm, dt)
    cd.Initialize()
    bgU = cd.advance(1, nt - 1, cols)
    dm.close()
    r = reduced(m, box, nt, bgU, lag=2, cols=cols)
    dr = DeviceModel(r)
    cr = CentralDifference(dr, dt)
    cr.Initialize()
    hist = cr.advance(1, nt - 1, cols)
    pos = np.full(m.nfree, -1)
    pos[cols] = np.arange(cols.size)
    ci = pos[check[np.isin(check, ins)]]
    co = pos[check[np.isin(check, out)]]
    scale = np.abs(bgU).max()
    assert np.abs(bgU[:, ci]).max() > 0.01 * scale and np.abs(bgU[:, co]).max() > 0.01 * scale
    assert np.abs(hist[:, ci] - bgU[:, ci]).max() <= 1e-9 * scale
    assert np.abs(hist[:, co]).max() <= 1e-9 * scale


def test_newton_j2_interior_matches_oracle():
    """NewmarkBeta + NewtonRaphson with J2 elements inside the box and an elastic DRM layer:
    the DRM forces enter every Newton residual."""
    from test_gpu_nonlinear import check
    m, r, box, dt, nt, rec, bg = case("quad", oracle.EXTENDED_NEWMARK)
    nt = 60
    x0, xl = (np.asarray(v, float) for v in box)
    centre = r.coords[r.conn].mean(axis=1)
    inner = np.nonzero(np.all(np.abs(centre - x0) < xl - 0.5, axis=1))[0]
    assert inner.size > 0 and not np.intersect1d(inner, r.drm_elem).size
    r.drm_signal = np.ascontiguousarray(r.drm_signal[:, :nt])
    r.set_j2(8.33e7, 3.85e7, 2e4, 4e6, 1.0, inner)
    check(r, dt, nt, list(range(r.nfree)), 1e-6, 20, 5, branch_free=False)


def test_batched_parts_with_own_dt_match_oracle():
    """Two DRM-loaded parts with their own time steps in one batched launch."""
    from svl_amd import batch as B
    nt = 120
    parts, dts, refs = [], [1e-3, 8e-4], []
    for i, dt in enumerate(dts):
        m, box = drm_quad_model(nt, dt)
        m.E = m.E * (1.0 - 0.2 * i)
        bg = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, list(range(m.nfree)))
        r = reduced(m, box, nt, bg["U"], lag=2)
        parts.append(r)
        refs.append(oracle.run(r, oracle.CENTRAL_DIFFERENCE, dt, nt, list(range(r.nfree)))["U"])
    u, nptr, fptr = B.union(parts)
    dm = DeviceModel(u)
    cd = B.BatchedCentralDifference(dm, dts, nptr, fptr)
    cd.Initialize()
    hist = cd.advance(1, nt - 1, rec_dofs=np.arange(u.nfree, dtype=np.int32))
    for i in range(len(parts)):
        assert rel_l2(hist[:, fptr[i]:fptr[i + 1]], refs[i]) <= 1e-10


def test_short_signal_and_factor_match_oracle():
    """Signals shorter than the run (rows past the end read as zero) and a load factor."""
    dt, nt = 1e-3, 200
    m, box = drm_quad_model(nt, dt)
    m.ray_am = np.full(m.nelem, 0.3)
    rec = list(range(m.nfree))
    bg = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    r = reduced(m, box, nt // 2, bg["U"], lag=2)
    r.drm_factor = 1.5
    assert r.drm_signal.shape[1] == nt // 2
    ref = oracle.run(r, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    dm = DeviceModel(r)
    cd = CentralDifference(dm, dt)
    cd.Initialize()
    hist = cd.advance(1, nt - 1, rec)
    assert np.abs(ref["U"][nt // 2:]).max() > 0
    assert rel_l2(hist, ref["U"]) <= 1e-10
