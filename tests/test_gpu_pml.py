"""2D PML on the device (svlg_pml_*) against the numpy oracle (oracle/pml2d.py).

* SVL's validation case F11 (10 PML2DQuad4 under 90 lin2DQuad4 elements, EQUAL constraints,
  dt = 0.01, its own Ricker series from tests/golden/F11.npz), all 999 steps: U, V, A of the
  interface, top and a PML stress unknown within 1e-10 relative L2 of the oracle's dense solve.
* A half-space wrapped by PML on three sides (corner elements stretched both ways), lumped and
  consistent regular mass, against the oracle.
* SVL's validation case J12 (10 PML3DHexa8 under 90 lin3DHexa8, EQUAL ties on uz), all 999
  steps, and a 3D half-space wrapped by PML on five faces, whose edge and corner elements
  carry the history matrix G (Ubar in Feff and the update), against the oracle.
* At a size the dense oracle does not reach, properties every linear time-invariant model
  holds: a load delayed by d steps gives the response delayed by d steps, and doubling the
  load doubles the response (to the MINRES tolerance).
"""
import time
from pathlib import Path

import numpy as np
import pytest

from _pml_meshes import rel_l2, to_oracle
from svl_amd.pml import PmlNewmark, halfspace_box, halfspace_box3d, soil_column, soil_rod3d

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
TOL = 1e-10


def ricker(nt, dt, t0=0.15, f0=10.0):
    a = (np.pi * f0 * (np.arange(nt) * dt - t0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


def reg_node(m, x, y):
    """The regular (2-unknown) node at (x, y)."""
    hit = np.nonzero((np.abs(m.coords - [x, y]).sum(axis=1) < 1e-9) & (m.ndof == 2))[0]
    assert hit.size == 1
    return int(hit[0])


def test_f11_column_matches_oracle():
    load = np.load(GOLD / "F11.npz")["load"]
    m = soil_column(90, 10, load)
    nt = load.size
    r0 = 2 * 11
    # xx and yy stress unknowns of a PML node (xy vanishes by symmetry in this column)
    rec = [(r0, 1), (len(m.ndof) - 2, 1), (10, 2), (10, 3), (r0 - 2, 1)]
    print("F11 column: 1000 steps, dt 0.01", flush=True)
    dev = PmlNewmark(m, 0.01)
    hu, hv, ha, its = dev.advance(1, nt - 1, rec)
    ou, ov, oa = to_oracle(m).run_newmark(0.01, nt, rec, k0=1, uva=True)
    for c in range(len(rec)):
        for h, o in ((hu, ou), (hv, ov), (ha, oa)):
            assert rel_l2(h[:, c], o[:, c]) < TOL, (rec[c], rel_l2(h[:, c], o[:, c]))
    # the tied interface moves with its master
    np.testing.assert_array_equal(hu[:, 4], hu[:, 0])
    assert its.max() < 4000 and its.min() >= 0
    print("MINRES iterations per step: mean", its.mean(), "max", its.max())


@pytest.mark.parametrize("lumped", [False, True])
def test_halfspace_box_matches_oracle(lumped):
    nt, dt = 121, 0.005
    m = halfspace_box(10, 6, 3, lumped=lumped, load_series=ricker(nt, dt))
    top = m.loads[0][0]
    off = reg_node(m, 3.0 + 2.0, 9.0)        # on the surface, 3 elements left of the load
    deep = reg_node(m, 3.0 + 8.0, 4.0)
    rec = [(top, 1), (off, 0), (off, 1), (deep, 0), (deep, 1)]
    # a corner PML node (all five unknowns) and a side interface node
    corner = int(np.argmin(np.sum((m.coords - [1.0, 1.0]) ** 2, axis=1)))
    assert m.ndof[corner] == 5
    rec += [(corner, a) for a in range(5)]
    s, _, mm, _, _ = m.equal[0]
    rec += [(s, 0), (mm, 0)]
    print(f"half-space box 16x9, lumped={lumped}, {nt - 1} steps", flush=True)
    dev = PmlNewmark(m, dt)
    hu, hv, ha, its = dev.advance(1, nt - 1, rec)
    ou, ov, oa = to_oracle(m).run_newmark(dt, nt, rec, k0=1, uva=True)
    for c in range(len(rec)):
        for h, o in ((hu, ou), (hv, ov), (ha, oa)):
            if np.linalg.norm(o[:, c]) == 0.0:
                assert np.abs(h[:, c]).max() == 0.0
            else:
                assert rel_l2(h[:, c], o[:, c]) < TOL, (rec[c], rel_l2(h[:, c], o[:, c]))
    U = dev.GetDisplacements()
    assert np.all(np.isfinite(U))


def test_large_box_time_invariance_and_linearity():
    nt, dt, d = 61, 0.005, 5
    s = ricker(nt, dt)
    # steps start at k = 1: the delayed run sees s[1], s[2], ... from step d + 1 on
    sd = np.concatenate([np.zeros(d + 1), s[1:nt - d]])
    m1 = halfspace_box(160, 80, 10, load_series=s)
    top = m1.loads[0][0]
    q = reg_node(m1, 10.0 + 60.0, 70.0)
    rec = [(top, 1), (q, 0), (q, 1)]
    print("half-space box 180x90 (10-element PML): delay and scaling", flush=True)
    a = PmlNewmark(m1, dt)
    print("info", a.info(), flush=True)
    t0 = time.perf_counter()
    ua, _, _, its = a.advance(1, nt - 1, rec)
    print(f"{(time.perf_counter() - t0) / (nt - 1) * 1e3:.2f} ms per step", flush=True)
    m2 = halfspace_box(160, 80, 10, load_series=sd)
    ub, _, _, _ = PmlNewmark(m2, dt).advance(1, nt - 1, rec)
    m3 = halfspace_box(160, 80, 10, load_series=2 * s)
    uc, _, _, _ = PmlNewmark(m3, dt).advance(1, nt - 1, rec)
    assert np.linalg.norm(ua) > 0
    assert rel_l2(ub[d:], ua[:-d]) < 1e-9
    assert np.abs(ub[:d]).max() <= 1e-9 * np.abs(ua).max()
    assert rel_l2(uc, 2 * ua) < 1e-9
    print("MINRES iterations per step: mean", its.mean(), "max", its.max())


def test_j12_rod_matches_oracle():
    load = np.load(GOLD / "F11.npz")["load"]    # J12 applies the same Ricker file as F11
    m = soil_rod3d(90, 10, load)
    nt = load.size
    r0 = 4 * 11
    top = len(m.ndof) - 4
    rec = [(r0, 2), (top, 2), (20, 5), (20, 3), (r0 - 4, 2)]
    print("J12 rod: 1000 steps, dt 0.01", flush=True)
    pm = PmlNewmark(m, 0.01)
    t0 = time.perf_counter()
    hu, hv, ha, its = pm.advance(1, nt - 1, rec)
    print(f"J12 rod: {(time.perf_counter() - t0) / (nt - 1) * 1e3:.3f} ms per step, {pm.info()}", flush=True)
    ou, ov, oa = to_oracle(m).run_newmark(0.01, nt, rec, k0=1, uva=True)
    for c in range(len(rec)):
        for h, o in ((hu, ou), (hv, ov), (ha, oa)):
            assert rel_l2(h[:, c], o[:, c]) < TOL, (rec[c], rel_l2(h[:, c], o[:, c]))
    np.testing.assert_array_equal(hu[:, 4], hu[:, 0])
    print("MINRES iterations per step: mean", its.mean(), "max", its.max())


@pytest.mark.parametrize("lumped", [False, True])
def test_halfspace_box3d_with_history_matrix(lumped):
    nt, dt = 81, 0.005
    m = halfspace_box3d(4, 4, 3, 1, lumped=lumped, load_series=ricker(nt, dt, t0=0.05, f0=20.0))
    top = m.loads[0][0]
    off = reg_node3(m, 2.0, 2.0, 4.0)        # off both symmetry planes
    corner = int(np.argmin(np.sum((m.coords - [0.0, 1.0, 1.0]) ** 2, axis=1)))
    assert m.ndof[corner] == 9
    # the six stress unknowns of a node on the outer face of a corner column (its
    # displacements are fixed)
    rec = [(top, 2), (off, 0), (off, 1), (off, 2)] + [(corner, a) for a in range(3, 9)]
    print(f"3D half-space box 6x6x4 with PML on five faces, lumped={lumped}, {nt - 1} steps", flush=True)
    pn = PmlNewmark(m, dt)
    hu, hv, ha, its = pn.advance(1, nt - 1, rec)
    orc = to_oracle(m)
    ou, ov, oa = orc.run_newmark(dt, nt, rec, k0=1, uva=True)
    for c in range(len(rec)):
        for h, o in ((hu, ou), (hv, ov), (ha, oa)):
            assert np.linalg.norm(o[:, c]) > 0
            assert rel_l2(h[:, c], o[:, c]) < TOL, (rec[c], rel_l2(h[:, c], o[:, c]))
    print("MINRES iterations per step: mean", its.mean(), "max", its.max())
    # GetPMLHistoryVector: the device's Ubar (free DOFs) against the oracle's after the last step,
    # at the recorded unknowns (the bar above); the whole-vector figures of Ubar and U are printed.
    off_t, T = orc.assemble()[:2]
    full = orc.run_newmark(dt, nt, rec, k0=1, full=True)
    ub_ref, u_ref = full[3][-1], full[0][-1]
    ub = T @ pn.GetPMLHistoryVector()
    u = T @ pn.GetDisplacements()
    idx = [off_t[n] + a for n, a in rec]
    print(f"Ubar at the recorded unknowns {rel_l2(ub[idx], ub_ref[idx]):.2e}; whole vectors: "
          f"Ubar {rel_l2(ub, ub_ref):.2e}, U {rel_l2(u, u_ref):.2e}")
    assert np.linalg.norm(ub_ref[idx]) > 0
    assert rel_l2(ub[idx], ub_ref[idx]) < TOL, rel_l2(ub[idx], ub_ref[idx])


def reg_node3(m, x, y, z):
    hit = np.nonzero((np.abs(m.coords - [x, y, z]).sum(axis=1) < 1e-9) & (m.ndof == 3))[0]
    assert hit.size == 1
    return int(hit[0])


@pytest.mark.parametrize("case", ["F11", "J12"])
def test_run_file_recorders(tmp_path, case):
    """SVL's F11 (Newmark) and J12 (ExtendedNewmark) written as run files and run by
    `python -m svl_amd.svlfile RUN.json`: the NODE recorder files of SVL's recorded nodes
    (F11: 23 and 203; J12: 45 and 405) against the oracle."""
    from svl_amd import svlfile
    load = np.load(GOLD / "F11.npz")["load"]
    if case == "F11":
        m, integ, tags = soil_column(90, 10, load), "NEWMARK", [23, 203]
    else:
        m, integ, tags = soil_rod3d(90, 10, load), "EXTENDEDNEWMARK", [45, 405]
    from svl_amd.pmlfile import write_pml_run
    path = str(tmp_path / "run.json")
    recs = [(f"{r}.out", r, tags, 15) for r in ("DISP", "VEL", "ACCEL")]
    write_pml_run(m, path, 0.01, load.size, integrator=integ, recorders=recs)
    print(f"{case} run file: {load.size - 1} steps", flush=True)
    svlfile.main([path, "--out", str(tmp_path / "out")])
    rec = [(t - 1, a) for t in tags for a in range(int(m.ndof[t - 1]))]
    ou, ov, oa = to_oracle(m).run_newmark(0.01, load.size, rec, k0=1, uva=True)
    for r, o in (("DISP", ou), ("VEL", ov), ("ACCEL", oa)):
        ftags, tot, ntotal, nt, vals = svlfile.read_node_recorder(str(tmp_path / "out" / f"{r}.out"))
        assert ftags == tags and nt == load.size and vals.shape == o.shape
        for c in range(o.shape[1]):
            if np.linalg.norm(o[:, c]) > 0:
                assert rel_l2(vals[:, c], o[:, c]) < TOL, (r, rec[c])


def test_minres_iteration_limit_stops():
    """A solve that cannot meet rtol within maxit stops with SVL's stop semantics (the limit is
    checked on the host between batches; every kernel exits normally)."""
    from svl_amd._capi import SvlgError
    m = soil_column(40, 10, ricker(20, 0.01))
    pm = PmlNewmark(m, 0.01, maxit=1)
    with pytest.raises(SvlgError, match="did not converge"):
        pm.advance(1, 5)


def test_eager_launches_equal_graph_replays():
    """svlg_pml_set_graph(0): the MINRES iterations launched eagerly give the bits of the
    graph-replayed ones (same kernels, same order) on a 3D box with G and Ubar."""
    nt, dt = 41, 0.005
    m = halfspace_box3d(3, 3, 2, 1, load_series=ricker(nt, dt, t0=0.05, f0=20.0))
    rec = [(m.loads[0][0], 2)]
    runs = []
    for graph in (True, False):
        pn = PmlNewmark(m, dt)
        pn.set_graph(graph)
        hu, hv, ha, its = pn.advance(1, nt - 1, rec)
        runs.append((hu, its, pn.GetPMLHistoryVector()))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][2], runs[1][2]) and np.abs(runs[0][2]).max() > 0
