"""GPU parity of support motion (Supports + POINTLOAD_SUPPORT_MOTION) against the oracle, whose
support-motion restatement is pinned by the relative-coordinate identities of
test_oracle_support_motion.

Models: tests/_meshes.support_model (Quad4 / Hexa8 blocks whose base nodes are restrained and
move in x, or in x and z, with a series cut short so that SVL's "first value past the end" rule
acts), with point loads, Rayleigh damping, Lysmer dashpots on the sides and consistent mass where
the integrator allows them. Bars: free-DOF history relative L2 <= 1e-10 (implicit: CG at 1e-13);
the total displacement vector after the last step (svlg_get_displacements: free DOFs and the
moving base) against the oracle's last row, <= 1e-10, the base DOFs bit for bit (the device
adds the same increments in the same order); reactions at the base <= 1e-9 / 1e-10; the graph
replay path bit-identical to the eager steps.
"""
import numpy as np
import pytest

import oracle
from svl_amd import _capi as capi
from svl_amd.integrators import (CentralDifference, DeviceModel, ExtendedNewmarkBeta, NewmarkBeta,
                                 total_displacements)

from _meshes import support_model

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def all_dofs(m, bnd):
    """Free DOFs, then (oracle convention -1 - t) the total DOFs of the base nodes."""
    nf = m.nfree
    base = [int(n) for n in bnd["bottom"]]
    tot = [-1 - (n * m.ndim + c) for n in base for c in range(m.ndim)]
    return list(range(nf)), tot, base


def check_total(dm, m, ref_last, base):
    """The device's total displacement against the oracle's last row (free, then base DOFs)."""
    Ut = total_displacements(dm)
    fd = m.free_dof()
    got_free = np.zeros(m.nfree)
    mask = fd >= 0
    got_free[fd[mask]] = Ut[mask]
    got_base = Ut[base].reshape(-1)
    nf = m.nfree
    e_free = rel(got_free, ref_last[:nf])
    print(f"  total displacement after the last step: free rel L2 {e_free:.2e}, base max |diff| "
          f"{np.max(np.abs(got_base - ref_last[nf:])):.1e}")
    assert e_free < 1e-10
    assert np.array_equal(got_base, ref_last[nf:])


@pytest.mark.parametrize("ndim,variant,am,dash,dirs", [
    (3, 3, 0.6, True, (0,)),
    (3, 3, 0.0, False, (0, 2)),
    (2, 3, 0.8, True, (0, 1)),
    (2, 2, 0.8, True, (0,)),
])
def test_central_difference_support_motion(ndim, variant, am, dash, dirs):
    dt, nt = 2e-4, 420
    m, g, bnd = support_model(ndim, nt, dt, am=am, dashpots=dash, length=300, factor=0.7, dirs=dirs)
    rec, tot, base = all_dofs(m, bnd)
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec + tot)["U"]
    dm = DeviceModel(m, variant=variant)
    cd = CentralDifference(dm, dt)
    cd.Initialize()
    got = cd.advance(1, nt - 1, rec)
    err = rel(got, ref[:, :len(rec)])
    print(f"CD {ndim}D variant {variant} am={am} dashpots={dash} dirs={dirs}: free history rel L2 {err:.2e}")
    assert err < 1e-10
    check_total(dm, m, ref[-1], base)
    dm.close()


def test_central_difference_graph_replay_matches_eager():
    """hipGraph replays (the support kernels take the step index from the device counter)."""
    dt, nt = 2e-4, 1 + 3 * 60
    m, g, bnd = support_model(3, nt, dt, am=0.4, dashpots=True, length=100, factor=1.1, dirs=(0, 1))
    rec, tot, base = all_dofs(m, bnd)
    out = []
    for graph in (0, 30):
        dm = DeviceModel(m, variant=3)
        cd = CentralDifference(dm, dt)
        cd.Initialize()
        if graph:
            capi.check(capi.lib().svlg_cd_graph_steps(dm.h, 1, nt - 1, graph), "svlg_cd_graph_steps")
        else:
            cd.advance(1, nt - 1)
        out.append(total_displacements(dm))
        dm.close()
    assert np.array_equal(out[0], out[1])
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec + tot)["U"][-1]
    fd = m.free_dof()
    mask = fd >= 0
    got = np.zeros(m.nfree)
    got[fd[mask]] = out[1][mask]
    err = rel(got, ref[:m.nfree])
    print(f"CD graph replay: bit-identical to eager; vs oracle rel L2 {err:.2e}")
    assert err < 1e-10


@pytest.mark.parametrize("integrator,offset,ndim,lumped,am,ak,dash", [
    (oracle.EXTENDED_NEWMARK, 0, 3, True, 0.5, 1e-3, True),
    (oracle.EXTENDED_NEWMARK, 0, 3, False, 0.5, 1e-3, True),
    (oracle.EXTENDED_NEWMARK, 0, 2, False, 0.0, 0.0, False),
    (oracle.NEWMARK, 1, 2, True, 0.4, 5e-4, True),
    (oracle.NEWMARK, 1, 3, False, 0.0, 2e-3, False),
])
def test_newmark_support_motion(integrator, offset, ndim, lumped, am, ak, dash):
    dt, nt = 1e-3, 140
    m, g, bnd = support_model(ndim, nt, dt, am=am, ak=ak, lumped=lumped, dashpots=dash, length=90, factor=1.2,
                              dirs=(0, ndim - 1))
    m.sm_offset = offset
    rec, tot, base = all_dofs(m, bnd)
    ref = oracle.run(m, integrator, dt, nt, rec + tot)["U"]
    dm = DeviceModel(m)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=5000)
    nm.Initialize()
    got = nm.advance(1, nt - 1, rec)
    err = rel(got, ref[:, :len(rec)])
    name = "ExtendedNewmarkBeta" if integrator == oracle.EXTENDED_NEWMARK else "NewmarkBeta (sample k-1)"
    print(f"{name} {ndim}D {'lumped' if lumped else 'consistent'} am={am} ak={ak} dashpots={dash}: "
          f"free history rel L2 {err:.2e}")
    assert err < 1e-10
    check_total(dm, m, ref[-1], base)
    dm.close()


def test_integrator_methods_with_support_motion():
    """ComputeEffectiveForce, ComputeSupportMotionVector, the solve and ComputeNewStep's update
    one by one (Linear.cpp:22-56), as a caller that keeps SVL's algorithm drives them."""
    dt, nt = 1e-3, 60
    m, g, bnd = support_model(3, nt, dt, am=0.5, ak=1e-3, lumped=False, dashpots=True, length=40)
    rec, tot, base = all_dofs(m, bnd)
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec + tot)["U"]
    dm = DeviceModel(m)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=5000)
    nm.Initialize()
    from svl_amd.integrators import LinearSystem
    ls = LinearSystem(nm)
    hist = []
    for k in range(1, nt):
        feff = nm.ComputeEffectiveForce(k)
        feff = nm.ComputeSupportMotionVector(feff, k)
        ls.SetVector(feff)
        ls.SolveSystem()
        nm.ApplyIncrement(ls.GetSolution(), k)
        hist.append(nm.GetDisplacements())
    err = rel(np.asarray(hist), ref[:, :len(rec)])
    print(f"ExtendedNewmarkBeta methods with support motion: rel L2 {err:.2e}")
    assert err < 1e-10
    check_total(dm, m, ref[-1], base)
    dm.close()


@pytest.mark.parametrize("integrator", [oracle.CENTRAL_DIFFERENCE, oracle.EXTENDED_NEWMARK])
def test_reactions_with_support_motion(integrator):
    """Reactions at the moving base: the element forces use the lagged displacement the strains
    hold (U_(k-1) at the support DOFs)."""
    dt, nt = (2e-4, 200) if integrator == oracle.CENTRAL_DIFFERENCE else (1e-3, 50)
    m, g, bnd = support_model(3, nt, dt, am=0.5, ak=0.0 if integrator == 0 else 1e-3, dashpots=True, length=120)
    nodes = np.asarray(bnd["bottom"], np.int32)
    tot_r = [int(n) * 3 + c for n in nodes for c in range(3)]
    ref = oracle.run(m, integrator, dt, nt, [], reac_tot=tot_r)["R"]
    dm = DeviceModel(m)
    integ = CentralDifference(dm, dt) if integrator == 0 else ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=5000)
    integ.Initialize()
    got = []
    for k in range(1, nt):
        integ.advance(k, 1)
        got.append(dm.reactions(nodes, k).reshape(-1))
    dm.close()
    err = rel(np.asarray(got), ref)
    print(f"reactions with support motion ({'CD' if integrator == 0 else 'ExtendedNewmarkBeta'}): rel L2 {err:.2e}")
    assert err < (1e-10 if integrator == 0 else 1e-9)


def test_newton_with_support_motion_refused():
    m, g, bnd = support_model(2, 20, 1e-3)
    m.set_j2(1e8, 4e7, 1e5, 1e6, 1.0)
    dm = DeviceModel(m)
    with pytest.raises(RuntimeError, match="Newton"):
        NewmarkBeta(dm, 1e-3).Initialize()
    dm.close()
