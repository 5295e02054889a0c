"""Pins the oracle's domain reduction load (lin*::ComputeDomainReductionForces with SVL's
exterior negation) by free-field reproduction.

A background run on a mesh with a source outside a box gives the free field. The same mesh
without the source, loaded by the DRM layer of the box with that free field as node signals,
must reproduce the free field at every node inside the box (boundary nodes included) and
stay at rest at every node outside it. This holds exactly for the discrete equations when the
forces are K_be u_e on the boundary side and K_eb u_b on the exterior side (plus the damping
and mass blocks): a flipped sign, a mask that keeps same-side blocks, a wrong motion channel
(U, V, A) or a shifted row breaks it; the planted variants below (flipped flags, negated signals,
shifted rows) show that the check sees them.

Row alignment: Newmark's step k balances forces at the new state (signal row k = state after
step k); CentralDifference's step k evaluates the stiffness at the state of step k - 1.
"""
import numpy as np
import pytest

import oracle
from svl_amd import drm
from svl_amd.mesh import Model

from _meshes import drm_hexa_model as hexa_model, drm_quad_model as quad_model


def reduced(m, box, sig_args, flip_cond=False, scale=1.0):
    elems, nodes, cond = drm.drm_box(m, *box)
    r = Model(**{k: getattr(m, k) for k in ("ndim", "coords", "node_tags", "elem_type", "conn", "E", "nu", "rho",
                                             "restrained", "ray_am", "ray_ak", "massform_lumped")})
    sig = drm.signals_from_history(m, nodes, *sig_args) * scale
    drm.attach(r, elems, nodes, (1 - cond) if flip_cond else cond, sig)
    return r, elems, nodes, cond


def reproduction_error(m, box, integ, dt, nt, lag, **plant):
    rec = list(range(m.nfree))
    bg = oracle.run(m, integ, dt, nt, rec, want_va=True)
    V, A = (bg["V"], bg["A"]) if integ != oracle.CENTRAL_DIFFERENCE else (None, None)
    r, elems, nodes, cond = reduced(m, box, (nt, bg["U"], V, A, lag), **plant)
    out = oracle.run(r, integ, dt, nt, rec)
    fd = m.free_dof()
    x0, xl = (np.asarray(v, float) for v in box)
    inside = np.all(np.abs(m.coords - x0) <= xl, axis=1)
    nd = m.ndim
    ins = np.array([fd[n, c] for n in np.nonzero(inside)[0] for c in range(nd) if fd[n, c] >= 0])
    out_dofs = np.array([fd[n, c] for n in np.nonzero(~inside)[0] for c in range(nd) if fd[n, c] >= 0])
    scale = np.abs(bg["U"]).max()
    assert np.abs(bg["U"][:, ins]).max() > 0.05 * scale, "the wave must reach the box"
    assert elems.size > 0 and cond.sum() > 0 and (1 - cond).sum() > 0
    e_in = np.abs(out["U"][:, ins] - bg["U"][:, ins]).max() / scale
    e_out = np.abs(out["U"][:, out_dofs]).max() / scale
    return e_in, e_out


CASES = [
    ("quad-cd", quad_model, oracle.CENTRAL_DIFFERENCE, 1e-3, 260, 2, True, 0.0, 0.0),
    ("quad-newmark-lumped", quad_model, oracle.EXTENDED_NEWMARK, 2e-3, 140, 1, True, 0.4, 3e-4),
    ("quad-newmark-consistent", quad_model, oracle.EXTENDED_NEWMARK, 2e-3, 140, 1, False, 0.4, 3e-4),
    ("hexa-newmark-lumped", hexa_model, oracle.EXTENDED_NEWMARK, 2e-3, 70, 1, True, 0.2, 2e-4),
]


@pytest.mark.parametrize("name,mk,integ,dt,nt,lag,lumped,am,ak", CASES, ids=[c[0] for c in CASES])
def test_free_field_reproduced(name, mk, integ, dt, nt, lag, lumped, am, ak):
    m, box = mk(nt, dt)
    m.massform_lumped = lumped
    m.ray_am = np.full(m.nelem, am)
    m.ray_ak = np.full(m.nelem, ak)
    e_in, e_out = reproduction_error(m, box, integ, dt, nt, lag)
    assert e_in < 1e-9 and e_out < 1e-9, (e_in, e_out)


@pytest.mark.parametrize("plant", ["flip_cond", "negate", "lag"])
def test_planted_mistakes_break_reproduction(plant):
    dt, nt = 2e-3, 140
    m, box = quad_model(nt, dt)
    m.massform_lumped = False
    m.ray_am = np.full(m.nelem, 0.4)
    m.ray_ak = np.full(m.nelem, 3e-4)
    kw = {"flip_cond": {"flip_cond": True}, "negate": {"scale": -1.0}, "lag": {}}[plant]
    lag = 2 if plant == "lag" else 1
    e_in, e_out = reproduction_error(m, box, oracle.EXTENDED_NEWMARK, dt, nt, lag, **kw)
    assert max(e_in, e_out) > 1e-2


def test_drm_box_matches_layer_definition():
    m, box = quad_model(10, 1e-3)
    elems, nodes, cond = drm.drm_box(m, *box)
    # A 7x5-node box in a unit grid: the layer is the ring of elements around the 6x4 inner
    # elements, (6+2)*(4+2) - 6*4 = 24 elements; nodes: 7*5 - 5*3 inner boundary ones (cond 0)
    # plus 9*7 - 7*5 outer ones (cond 1).
    assert elems.size == 24
    assert int((cond == 0).sum()) == 7 * 5 - 5 * 3
    assert int((cond == 1).sum()) == 9 * 7 - 7 * 5
    assert np.all(np.diff(nodes) > 0) and np.all(np.diff(elems) > 0)
