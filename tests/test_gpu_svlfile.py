"""GPU: SVL run files through the device path, NODE recorders out (svl_amd.svlfile.run).

A model is written as an SVL run file, read back, run on the device by the integrator and
algorithm the file names, and its NODE recorder files (DISP, VEL, ACCEL) are parsed and
compared with the oracle's histories of the original model: relative L2 <= 1e-10 for DISP
(CG at 1e-13 on the implicit paths), <= 1e-8 for VEL / ACCEL (formed from the displacement
history by the integrator's update formulas, so differences of nearly equal displacements
enter), <= 1e-9 for REACTION at fixed mesh nodes and dashpot partner nodes, restrained
components of the motion exactly 0.
"""
import numpy as np
import pytest

import oracle
from svl_amd import svlfile as S
from svl_amd import workloads as W

from _meshes import block_model

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def case(kind):
    if kind == "cd":
        m, dt, nt = W.layer2d(16, 8, nt=200)
        return m, dt, 200, "CENTRALDIFFERENCE", "LINEAR", oracle.CENTRAL_DIFFERENCE, {}
    if kind in ("newmark", "newmark_default_mass"):
        m, dt, nt, src = block_model(4, 4, 3)
        m.ray_am = np.full(m.nelem, 0.5)
        m.ray_ak = np.full(m.nelem, 1e-3)
        m.massform_lumped = kind == "newmark"
        return m, dt, 120, "EXTENDEDNEWMARK", "LINEAR", oracle.EXTENDED_NEWMARK, {}
    m, dt, nt = W.soil_structure_j2(8, 8, 4, nt=41, block=(2, 2, 2), Ap=2e7, t0=0.08, f0=8.0)
    return m, dt, 41, "NEWMARK", "NEWTON", oracle.NEWMARK, dict(algorithm=oracle.NEWTON, nmax=20, cnvgtest=5,
                                                                   cnvgtol=1e-6)


@pytest.mark.parametrize("kind", ["cd", "newmark", "newmark_default_mass", "newton"])
def test_run_file_through_device(tmp_path, kind):
    """newmark_default_mass: the run file names no massform, so SVL's default (CONSISTENT,
    Driver.hpp:1998) applies."""
    m, dt, nt, integ, alg, ointeg, okw = case(kind)
    nodes = [0, m.nnodes // 3, m.nnodes - 1, int(m.load_node[0])]
    ntags = [n + 1 for n in nodes]
    recs = [("U.out", "DISP", ntags, 12), ("V.out", "VEL", ntags, 12), ("A.out", "ACCEL", ntags, 12)]
    # Reactions: fixed mesh nodes, and the partner nodes of the first dashpot nodes.
    fixed_nodes = list(np.nonzero(m.restrained.any(1))[0][:6])
    order = []
    for n in m.dash_node:
        if int(n) not in order:
            order.append(int(n))
    partner_idx = list(range(min(4, len(order))))
    rtags = [int(n) + 1 for n in fixed_nodes] + [m.nnodes + j + 1 for j in partner_idx]
    if rtags:
        recs.append(("R.out", "REACTION", rtags, 12))
    p = str(tmp_path / "run" / "model.json")
    import os
    os.makedirs(os.path.dirname(p))
    S.write_run(m, p, integ, dt, nt, alg, cnvgtol=1e-6, nmax=20, cnvgtest=5, recorders=recs)
    if kind == "newmark_default_mass":
        import json
        with open(p) as f:
            js = json.load(f)
        assert js["Global"].pop("massform") == "CONSISTENT"
        with open(p, "w") as f:
            json.dump(js, f)
    r = S.read_run(p)
    out_dir = str(tmp_path / "Solution")
    S.run(r, out_dir=out_dir, cg_rtol=1e-13)
    fd = m.free_dof()
    cols = [int(fd[n][c]) for n in nodes for c in range(m.ndim)]
    free_cols = [f for f in cols if f >= 0]
    reac_tot = ([int(n) * m.ndim + c for n in fixed_nodes for c in range(m.ndim)]
                + [(m.nnodes + j) * m.ndim + c for j in partner_idx for c in range(m.ndim)])
    ref = oracle.run(m, ointeg, dt, nt, free_cols, want_va=True, reac_tot=reac_tot, **okw)
    if rtags:
        tags, tot, ntotal, ntr, vals = S.read_node_recorder(os.path.join(out_dir, "R.out"))
        assert tags == rtags and vals.shape == (nt - 1, len(reac_tot))
        err = rel(vals, ref["R"])
        print(f"{kind} R.out: rel L2 {err:.2e}")
        assert err < 1e-9
    for fname, key, tol in (("U.out", "U", 1e-10), ("V.out", "V", 1e-8), ("A.out", "A", 1e-8)):
        tags, tot, ntotal, ntr, vals = S.read_node_recorder(os.path.join(out_dir, fname))
        assert tags == ntags and ntr == nt and vals.shape == (nt - 1, len(cols))
        restrained = [i for i, f in enumerate(cols) if f < 0]
        assert np.all(vals[:, restrained] == 0.0)
        free = [i for i, f in enumerate(cols) if f >= 0]
        err = rel(vals[:, free], ref[key])
        print(f"{kind} {fname}: rel L2 {err:.2e}")
        assert err < tol


def test_drm_run_file_through_device(tmp_path):
    """A DRM-loaded model (ElementLoad GENERALWAVE, one signal file per node) written, read
    and run on the device: DISP against the oracle's run of the original model."""
    import os
    from svl_amd import drm
    from _meshes import drm_quad_model
    dt, nt = 2e-3, 60
    m, box = drm_quad_model(nt, dt)
    m.ray_am = np.full(m.nelem, 0.4)
    m.ray_ak = np.full(m.nelem, 3e-4)
    bg = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, list(range(m.nfree)), want_va=True)
    m.load_node = m.load_node[:0]
    m.load_dir, m.load_factor, m.load_series = np.zeros((0, 2)), np.zeros(0), np.zeros((0, 1))
    e, n, c = drm.drm_box(m, *box)
    drm.attach(m, e, n, c, drm.signals_from_history(m, n, nt, bg["U"], bg["V"], bg["A"], 1))
    nodes = [int(n[0]), int(n[-1]), m.nnodes // 2]
    ntags = [k + 1 for k in nodes]
    p = str(tmp_path / "run" / "model.json")
    os.makedirs(os.path.dirname(p))
    S.write_run(m, p, "EXTENDEDNEWMARK", dt, nt, recorders=[("U.out", "DISP", ntags, 14)])
    out_dir = str(tmp_path / "Solution")
    S.run(S.read_run(p), out_dir=out_dir, cg_rtol=1e-13)
    fd = m.free_dof()
    cols = [int(fd[k][d]) for k in nodes for d in range(2)]
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, [f for f in cols if f >= 0])
    tags, tot, ntotal, ntr, vals = S.read_node_recorder(os.path.join(out_dir, "U.out"))
    free = [i for i, f in enumerate(cols) if f >= 0]
    assert np.abs(ref["U"]).max() > 0
    assert rel(vals[:, free], ref["U"]) < 1e-10


@pytest.mark.parametrize("integ", ["CENTRALDIFFERENCE", "EXTENDEDNEWMARK", "NEWMARK"])
def test_support_motion_run_file_through_device(tmp_path, integ):
    """A run file with Supports (TIMESERIES per DOF) and two SupportMotion loads: the DISP
    recorder of a base node holds its prescribed motion and a free node matches the oracle; the
    VEL recorder follows the integrator (Newmark: 0 at the support DOFs)."""
    from _meshes import support_model
    dt, nt = (2e-4, 240) if integ == "CENTRALDIFFERENCE" else (1e-3, 80)
    m, g, bnd = support_model(3, nt, dt, am=0.4, dashpots=True, length=nt // 2, factor=0.8, dirs=(0, 2))
    for n in bnd["bottom"][:4]:
        m.add_support_motion(int(n), 0, g, 0.3)
        m.add_support_motion(int(n), 2, g, 0.3)
    b0 = int(bnd["bottom"][1])
    top = int(bnd["top"][3])
    p = str(tmp_path / "sm.json")
    tags = S.write_run(m, p, integ, dt, nt, recorders=[("U.out", "DISP", [b0 + 1, top + 1], 10),
                                                       ("V.out", "VEL", [b0 + 1], 10)])
    r = S.read_run(p)
    out = S.run(r, out_dir=str(tmp_path), json_path=p, cg_rtol=1e-13)
    U = out[str(tmp_path / "U.out")]
    Vb = out[str(tmp_path / "V.out")]
    fd = m.free_dof()
    rec = [-1 - (b0 * 3 + c) for c in range(3)] + [int(fd[top, c]) for c in range(3)]
    io = {"CENTRALDIFFERENCE": oracle.CENTRAL_DIFFERENCE, "EXTENDEDNEWMARK": oracle.EXTENDED_NEWMARK,
          "NEWMARK": oracle.NEWMARK}[integ]
    ref = oracle.run(m, io, dt, nt, rec, want_va=True)
    dev = np.concatenate([U[:, 0, :], U[:, 1, :]], axis=1)
    err = rel(dev, ref["U"])
    print(f"{integ} run file with support motion: base + free node DISP vs oracle {err:.2e}, "
          f"base x at the end {U[-1, 0, 0]:.4e}")
    assert err < 1e-10
    assert np.array_equal(U[:, 0, :], ref["U"][:, :3])
    if integ == "CENTRALDIFFERENCE":
        assert rel(Vb[:, 0, :], ref["V"][:, :3]) < 1e-10
    else:
        assert np.all(Vb[:, 0, :] == 0.0)
