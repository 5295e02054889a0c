"""GPU: the run files SVL's own pre-processor writes for its validation cases
(tests/golden/svl_runfiles/, made by tests/golden/make_svl_runfiles.py with SVL's
CreateRunAnalysisFiles) read unchanged by svl_amd.svlfile / svl_amd.pmlfile and run on the
device through the run-file path (NODE recorders out, as SVL's driver writes them).

Bars: the recorded displacement history against the oracle's run of the model as read,
relative L2 <= 1e-10 (CG at 1e-13; the Newton cases' round-off branch of the J2 return map is
handled as in test_gpu_nonlinear: F07 at 10x its Newton tolerance); and against the OpenSees
histories SVL ships for the same cases (6 significant digits), at the bars the oracle's own
fixture tests meet: 1e-5 relative RMS for the elastic cases, 2e-3 for F03 and 1e-5 for F07.
F11 (PML) ships no reference history: the device against the oracle's PML restatement only.
"""
import os
from pathlib import Path

import numpy as np
import pytest

import oracle
from svl_amd import pmlfile as P
from svl_amd import svlfile as S

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
RUN = GOLD / "svl_runfiles"

# case: (recorded node tag, OpenSees columns of that node's x (and y / z) displacement,
#        OpenSees bar, oracle bar, steps kept)
CASES = {
    "F02": ("DISP", [1, 2], 1e-5, 1e-10, None),
    "J02": ("DISP", [3, 1, 2], 1e-5, 1e-10, None),
    "F06": ("DISP", None, 1e-5, 1e-10, None),
    "J05": ("DISP", None, 1e-5, 1e-10, None),
    "F03": ("DISP", [1], 2e-3, 1e-10, None),
    "F07": ("DISP", None, 1e-5, 1e-5, 2001),
}


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def run_file(tag):
    files = sorted((RUN / tag / "Partition").glob("*.json"))
    assert len(files) == 1
    return str(files[0])


@pytest.mark.parametrize("tag", list(CASES))
def test_svl_run_file_on_device(tmp_path, tag):
    resp, cols, bar_os, bar_or, keep = CASES[tag]
    path = run_file(tag)
    r = S.read_run(path)
    if keep:
        r.nt = keep
    rec = [x for x in r.recorders if x.response == resp][0]
    out = S.run(r, out_dir=str(tmp_path), json_path=path, cg_rtol=1e-13)
    vals = out[os.path.join(str(tmp_path), rec.file)]            # [steps, nodes, ndim]
    m = r.model
    tags = list(r.node_tags)
    fd = m.free_dof()
    nodes = [tags.index(t) for t in rec.tags]
    comp = [(i, c) for i, n in enumerate(nodes) for c in range(m.ndim) if fd[n][c] >= 0]
    dofs = [int(fd[nodes[i]][c]) for i, c in comp]
    dev = np.stack([vals[:, i, c] for i, c in comp], axis=1)
    kw = {}
    integ = oracle.NEWMARK
    if r.algorithm == "NEWTON":
        kw = dict(algorithm=oracle.NEWTON, nmax=r.nmax, cnvgtest=r.cnvgtest, cnvgtol=r.cnvgtol)
    ref = oracle.run(m, integ, r.dt, r.nt, dofs, **kw)["U"]
    e_or = rel(dev, ref)
    g = np.load(GOLD / f"{tag}.npz")
    osd = g["displacement"][:r.nt - 1]
    if cols is None:            # soil columns: x at the recorded nodes (base, top) in OpenSees columns 1, 3
        x_nodes = [j for j, (i, c) in enumerate(comp) if c == 0]
        os_cols = [1, 3][:len(x_nodes)] if len(x_nodes) == 2 else [3]
        dev_os = dev[:, x_nodes[-len(os_cols):]]
        os_u = osd[:, os_cols]
    else:
        dev_os = dev[:, :len(cols)]
        os_u = osd[:, cols]
    e_os = rel_rms(dev_os, os_u)
    print(f"{tag} run file (SVL pre-processor): {r.integrator}/{r.algorithm}, {r.nt - 1} steps, massform "
          f"{'LUMPED' if m.massform_lumped else 'CONSISTENT'}: vs oracle {e_or:.2e}, vs OpenSees {e_os:.2e}")
    assert e_or < bar_or
    assert e_os < bar_os


def test_svl_pml_run_file_on_device(tmp_path):
    """F11: 10 PML2DQuad4 + 90 lin2DQuad4 elements with EQUAL ties, as SVL writes it."""
    from _pml_meshes import rel_l2, to_oracle
    path = run_file("F11")
    pr = P.read_pml_run(path)
    out = P.run_pml(pr, out_dir=str(tmp_path), json_path=path)
    rec = [x for x in pr.recorders if x.response == "DISP"][0]
    vals = out[os.path.join(str(tmp_path), rec.file)]
    tags = pr.node_tags
    pairs = [(tags.index(t), a) for t in rec.tags for a in range(int(pr.mesh.ndof[tags.index(t)]))]
    ou = to_oracle(pr.mesh).run_newmark(pr.dt, pr.nt, pairs, k0=1)
    keep = [j for j in range(len(pairs)) if np.linalg.norm(ou[:, j]) > 0]
    err = rel_l2(vals[:, keep], ou[:, keep])
    print(f"F11 run file (SVL pre-processor): {pr.nt - 1} steps, {len(keep)} recorded unknowns in motion: "
          f"vs oracle {err:.2e}")
    assert keep and err < 1e-10
