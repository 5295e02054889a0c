"""SVL run-file reader / writer and NODE recorder format (CPU).

* write_run -> read_run reproduces the model (geometry, materials incl. J2, Rayleigh groups,
  Lysmer dashpots through restrained partner nodes, point loads with their series files and
  combination factors, the free-DOF numbering) and the simulation parameters;
* a run file written by hand in the schema SVL's pre-processor emits (upper-case names, node
  lists shared by one load, repeated nodes, a CONSTANT load) is read with SVL's semantics;
* the oracle gives the same history for the model read back as for the original;
* what is not on the device path is refused;
* NODE recorder files follow SVL's header / row layout and read back.
"""
import json
import os

import numpy as np
import pytest

import oracle
from svl_amd import svlfile as S
from svl_amd import workloads as W
from svl_amd.mesh import MAT_J2

from _meshes import block_model


def models():
    m1, dt1, nt1 = W.layer2d(12, 6, nt=60)
    m2, dt2, nt2, src = block_model(4, 3, 3)
    m2.ray_am = np.full(m2.nelem, 0.3)
    m2.ray_ak = np.where(np.arange(m2.nelem) % 2 == 0, 1e-3, 2e-3)
    m2.set_j2(1.3e8, 8e7, 4e5, 8e6, 0.5, np.arange(0, m2.nelem, 3))
    return [(m1, dt1, 60, "CENTRALDIFFERENCE", "LINEAR"), (m2, dt2, 40, "NEWMARK", "NEWTON")]


@pytest.mark.parametrize("case", [0, 1])
def test_write_read_roundtrip(tmp_path, case):
    m, dt, nt, integ, alg = models()[case]
    p = str(tmp_path / "model.json")
    S.write_run(m, p, integ, dt, nt, alg, cnvgtol=1e-7, nmax=17, cnvgtest=5,
                recorders=[("U.out", "DISP", [1, 4], 6)])
    r = S.read_run(p)
    q = r.model
    assert (r.integrator, r.algorithm, r.dt, r.nt, r.nmax, r.cnvgtest, r.cnvgtol) == (integ, alg, dt, nt, 17, 5, 1e-7)
    assert np.array_equal(q.coords, m.coords) and np.array_equal(q.conn, m.conn)
    assert np.array_equal(q.free_dof(), m.free_dof())
    assert np.allclose(q.E, m.E, rtol=1e-15) and np.allclose(q.nu, m.nu, rtol=1e-15)
    assert np.array_equal(q.rho, m.rho) and np.array_equal(q.mat_type, m.mat_type)
    assert np.array_equal(q.j2[q.mat_type == MAT_J2], m.j2[m.mat_type == MAT_J2])
    am = np.zeros(m.nelem) if m.ray_am is None else m.ray_am
    assert np.array_equal(q.ray_am, am)
    assert np.array_equal(q.dash_node, m.dash_node) and np.array_equal(q.dash_dir, m.dash_dir)
    assert np.array_equal(q.dash_eta, m.dash_eta)
    assert np.array_equal(q.load_node, m.load_node) and np.array_equal(q.load_factor, m.load_factor)
    assert np.array_equal(q.load_series, m.load_series)
    assert q.massform_lumped == m.massform_lumped and q.thickness == m.thickness
    assert r.recorders[0].tags == [1, 4] and r.recorders[0].ndps == 6
    # Same physics through the oracle.
    rec = list(range(0, m.nfree, 5))
    kw = dict(algorithm=oracle.NEWTON, nmax=17, cnvgtest=5, cnvgtol=1e-7) if alg == "NEWTON" else {}
    integ_o = oracle.CENTRAL_DIFFERENCE if integ == "CENTRALDIFFERENCE" else oracle.NEWMARK
    a = oracle.run(m, integ_o, dt, 12, rec, **kw)["U"]
    b = oracle.run(q, integ_o, dt, 12, rec, **kw)["U"]
    assert np.array_equal(a, b)


def handwritten(tmp_path, extra=None):
    ser = tmp_path / "ricker.txt"
    ser.write_text("5\n0.0 1.0 2.0 -1.0 0.5\n")
    js = {
        "Global": {"ndim": 2, "ntotal": 12, "nfree": 5, "massform": "LUMPED"},
        "Nodes": {
            "1": {"ndof": 2, "freedof": [-1, -1], "totaldof": [0, 1], "coords": [0.0, 0.0]},
            "2": {"ndof": 2, "freedof": [0, -1], "totaldof": [2, 3], "coords": [1.0, 0.0]},
            "3": {"ndof": 2, "freedof": [1, 2], "totaldof": [4, 5], "coords": [1.0, 1.0]},
            "4": {"ndof": 2, "freedof": [3, 4], "totaldof": [6, 7], "coords": [0.0, 1.0]},
            "5": {"ndof": 2, "freedof": [-1, -1], "totaldof": [8, 9], "coords": [1.0, 0.0]},
        },
        "Materials": {"1": {"name": "ELASTIC2DPLANESTRAIN", "attributes": {"E": 2e8, "nu": 0.25, "rho": 2000.0}},
                      "2": {"name": "VISCOUS1DLINEAR", "attributes": {"eta": 3e5}}},
        "Elements": {"1": {"name": "LIN2DQUAD4", "conn": [1, 2, 3, 4],
                           "attributes": {"material": 1, "th": 1.0, "rule": "GAUSS", "np": 4}},
                     "2": {"name": "ZEROLENGTH1D", "conn": [2, 5], "attributes": {"material": 2, "dir": 0}}},
        "Dampings": {"1": {"name": "Rayleigh", "attributes": {"am": 0.1, "ak": 0.0, "list": [1]}}},
        "Loads": {"1": {"name": "PointLoad", "attributes": {"type": "CONCENTRATED", "name": "TIMESERIES",
                                                            "file": str(ser), "list": [3, 4, 3], "dir": [0.0, -1.0]}},
                  "2": {"name": "PointLoad", "attributes": {"type": "CONCENTRATED", "name": "CONSTANT", "mag": 7.0,
                                                            "list": [4], "dir": [1.0, 0.0]}}},
        "Combinations": {"1": {"name": "Combo", "attributes": {"load": [1, 2], "factor": [2.0, 0.5]}}},
        "Simulations": {"combo": 1, "attributes": {
            "solver": {"name": "EIGEN", "update": "OFF"}, "algorithm": {"name": "LINEAR"},
            "integrator": {"name": "CentralDifference", "dt": 1e-3}, "analysis": {"name": "Dynamic", "nt": 5}}},
        "Recorders": {"1": {"name": "NODE", "file": "U.out", "resp": "disp", "ndps": 4, "nsamp": 1, "list": [3]},
                      "2": {"name": "PARAVIEW", "file": "pv", "features": 0, "ndps": 4, "nsamp": 10}},
    }
    if extra:
        extra(js)
    p = tmp_path / "hand.json"
    p.write_text(json.dumps(js))
    return str(p)


def test_handwritten_run_file(tmp_path):
    r = S.read_run(handwritten(tmp_path))
    m = r.model
    assert m.nnodes == 4 and m.nelem == 1 and m.nfree == 5
    assert np.array_equal(m.free_dof(), [[-1, -1], [0, -1], [1, 2], [3, 4]])
    assert list(m.dash_node) == [1] and list(m.dash_dir) == [0] and list(m.dash_eta) == [3e5]
    # Node 3 repeated in the list of load 1 counts once; load 2 is constant over the steps.
    assert list(m.load_node) == [2, 3, 3] and list(m.load_factor) == [2.0, 2.0, 0.5]
    assert np.array_equal(m.load_series[0], [0.0, 1.0, 2.0, -1.0, 0.5])
    assert np.array_equal(m.load_series[2], [7.0] * 5)
    assert r.integrator == "CENTRALDIFFERENCE" and r.nt == 5 and r.ntotal == 12
    assert [x.tags for x in r.recorders] == [[3]] and len(r.skipped_recorders) == 1
    assert np.array_equal(m.ray_am, [0.1])


@pytest.mark.parametrize("edit", [
    # Support motion on a free DOF (node 2, x), and with the Newton algorithm.
    lambda js: (js.update({"Supports": {"2": {"type": "Constant", "dof": [0], "value": [0.1]}}}),
                js["Loads"].update({"3": {"name": "SupportMotion", "attributes": {"list": [2]}}}),
                js["Combinations"]["1"]["attributes"].update({"load": [1, 2, 3], "factor": [2.0, 0.5, 1.0]})),
    lambda js: (js.update({"Supports": {"1": {"type": "Constant", "dof": [1], "value": [0.1]}}}),
                js["Loads"].update({"3": {"name": "SupportMotion", "attributes": {"list": [1]}}}),
                js["Combinations"]["1"]["attributes"].update({"load": [1, 2, 3], "factor": [2.0, 0.5, 1.0]}),
                js["Simulations"]["attributes"].update({"algorithm": {"name": "NEWTON"},
                                                        "integrator": {"name": "NEWMARK", "dt": 1e-3}})),
    lambda js: js["Elements"].update({"3": {"name": "lin2DQuad8", "conn": [1, 2, 3, 4], "attributes": {}}}),
    lambda js: js["Materials"]["1"].update({"name": "PlasticPlaneStrainBA"}),
    lambda js: js["Loads"]["1"]["attributes"].update({"type": "BODY"}),
    lambda js: js["Elements"]["1"]["attributes"].update({"np": 9}),
])
def test_refuses_what_is_not_on_the_path(tmp_path, edit):
    with pytest.raises(ValueError):
        S.read_run(handwritten(tmp_path, edit))


def test_handwritten_support_motion(tmp_path):
    """Supports as SVL's pre-processor writes them (dof 0-based; CONSTANT value / TIMESERIES file
    with the sample count first) activated by a SupportMotion load of the combination."""
    ser = tmp_path / "base.txt"
    ser.write_text("4\n0.1 0.2 0.3 0.4\n")

    def edit(js, integ="CentralDifference"):
        js["Supports"] = {"1": {"type": "TIMESERIES", "file": [str(ser), str(ser)], "dof": [1, 0]},
                          "2": {"type": "CONSTANT", "value": [0.05], "dof": [1]}}
        js["Loads"]["3"] = {"name": "SupportMotion", "attributes": {"list": [1, 2]}}
        js["Combinations"]["1"]["attributes"].update({"load": [1, 2, 3], "factor": [2.0, 0.5, 1.5]})
        js["Simulations"]["attributes"]["integrator"]["name"] = integ

    m = S.read_run(handwritten(tmp_path, edit)).model
    assert list(m.sm_node) == [0, 0, 1] and list(m.sm_dir) == [0, 1, 1]
    assert list(m.sm_factor) == [1.5] * 3 and m.sm_offset == 0
    assert np.array_equal(m.sm_series[0], [0.1, 0.2, 0.3, 0.4]) and np.array_equal(m.sm_series[2], [0.05])
    m2 = S.read_run(handwritten(tmp_path, lambda js: edit(js, "Newmark"))).model
    assert m2.sm_offset == 1
    m3 = S.read_run(handwritten(tmp_path, lambda js: edit(js, "ExtendedNewmark"))).model
    assert m3.sm_offset == 0


@pytest.mark.parametrize("integ", ["CENTRALDIFFERENCE", "EXTENDEDNEWMARK", "NEWMARK"])
def test_support_motion_roundtrip(tmp_path, integ):
    from _meshes import support_model
    dt = 2e-4 if integ == "CENTRALDIFFERENCE" else 1e-3
    m, g, bnd = support_model(2, 40, dt, am=0.3, dashpots=True, length=25, factor=0.9, dirs=(0, 1))
    for n in bnd["bottom"][:3]:
        m.add_support_motion(int(n), 0, g, 0.4)
        m.add_support_motion(int(n), 1, g, 0.4)
    p = str(tmp_path / "sm.json")
    S.write_run(m, p, integ, dt, 40)
    q = S.read_run(p).model
    key = lambda mm: sorted(zip(mm.sm_node.tolist(), mm.sm_dir.tolist(), mm.sm_factor.tolist(),  # noqa: E731
                                [tuple(x) for x in mm.sm_series]))
    assert key(q) == key(m)
    assert q.sm_offset == (1 if integ == "NEWMARK" else 0)
    io = {"CENTRALDIFFERENCE": oracle.CENTRAL_DIFFERENCE, "EXTENDEDNEWMARK": oracle.EXTENDED_NEWMARK,
          "NEWMARK": oracle.NEWMARK}[integ]
    rec = list(range(0, m.nfree, 3)) + [-1 - int(bnd["bottom"][1]) * 2]
    a = oracle.run(m, io, dt, 30, rec)["U"]
    b = oracle.run(q, io, dt, 30, rec)["U"]
    assert np.array_equal(a, b) and np.abs(a[:, -1]).max() > 0


def test_node_recorder_format(tmp_path):
    rec = S.NodeRecorder("X.out", "DISP", [7, 3], 5)
    vals = np.array([[[0.125, -3.5], [1e-3, 0.0]], [[2.0, 4.0], [-0.25, 1e10]], [[1.0, 1.0], [1.0, 1.0]]])
    p = str(tmp_path / "X.out")
    S.write_node_recorder(p, rec, [3, 7], np.array([[0, 1], [2, 3]]), 9, 4, vals)
    lines = open(p).read().splitlines()
    assert lines[0] == "2 4 9 4"
    assert lines[1] == "7 2 2 3" and lines[2] == "3 2 0 1"
    assert lines[3] == "1.25000e-01 -3.50000e+00 1.00000e-03 0.00000e+00 "
    assert lines[4] == "2.00000e+00 4.00000e+00 -2.50000e-01 1.00000e+10 "
    tags, tot, ntotal, nt, v = S.read_node_recorder(p)
    assert tags == [7, 3] and ntotal == 9 and nt == 4
    assert np.allclose(v, vals.reshape(3, 4), rtol=1e-5, atol=1e-9)


def test_drm_roundtrip(tmp_path):
    from svl_amd import drm
    from _meshes import drm_quad_model
    dt, nt = 2e-3, 30
    m, box = drm_quad_model(nt, dt)
    rec = list(range(m.nfree))
    bg = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec, want_va=True)
    m.load_node = m.load_node[:0]
    m.load_dir, m.load_factor, m.load_series = np.zeros((0, 2)), np.zeros(0), np.zeros((0, 1))
    elems, nodes, cond = drm.drm_box(m, *box)
    elems = elems[::-1].copy()  # SVL keeps the list order of the load's elements
    drm.attach(m, elems, nodes, cond, drm.signals_from_history(m, nodes, nt, bg["U"], bg["V"], bg["A"], 1), 0.75)
    p = str(tmp_path / "drm.json")
    S.write_run(m, p, "EXTENDEDNEWMARKBETA", dt, nt)
    js = json.load(open(p))
    ld = js["Loads"]["1"]
    assert ld["name"] == "ElementLoad" and ld["attributes"]["type"] == "GENERALWAVE"
    assert js["Combinations"]["1"]["attributes"]["factor"] == [0.75]
    q = S.read_run(p).model
    assert np.array_equal(q.drm_elem, elems) and np.array_equal(q.drm_node, nodes)
    assert np.array_equal(q.drm_cond, cond) and np.array_equal(q.drm_signal, m.drm_signal)
    assert q.drm_factor == 0.75
    a = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, rec)["U"]
    b = oracle.run(q, oracle.EXTENDED_NEWMARK, dt, nt, rec)["U"]
    assert np.abs(a).max() > 0 and np.array_equal(a, b)


def test_handwritten_generalwave(tmp_path):
    # Node files as SVL's driver reads them: "nt nFields cond", then nt rows.
    for tag, c in ((1, 1), (2, 0), (3, 0), (4, 1)):
        rows = [[tag + 0.1 * k + 0.01 * f for f in range(6)] for k in range(3)]
        (tmp_path / f"drm{tag}.txt").write_text(f"3 6 {c}\n" + "\n".join(" ".join(map(str, r)) for r in rows))

    def edit(js):
        js["Loads"]["3"] = {"name": "ElementLoad", "attributes": {"type": "GENERALWAVE", "name": "TIMESERIES",
                                                                  "file": str(tmp_path / "drm$.txt"), "list": [1]}}
        js["Combinations"]["1"]["attributes"] = {"load": [1, 3], "factor": [2.0, 1.5]}
    m = S.read_run(handwritten(tmp_path, edit)).model
    assert list(m.drm_elem) == [0] and list(m.drm_node) == [0, 1, 2, 3] and list(m.drm_cond) == [1, 0, 0, 1]
    assert m.drm_factor == 1.5 and m.drm_signal.shape == (4, 3, 6)
    assert m.drm_signal[3, 2, 5] == 4 + 0.2 + 0.05  # stored as read; the exterior sign flips on the device


def test_reads_run_files_written_by_svl_preprocessor():
    """The run files SVL's own pre-processor writes for its validation cases
    (tests/golden/svl_runfiles/, CreateRunAnalysisFiles output) are read unchanged: the free and
    total DOF counts, the mass form, the integrator and the load series agree with the JSON and
    the series files it names (looked up next to the run file, then in the case folder above
    Partition/, where SVL's driver finds them)."""
    import glob
    import json
    import os
    from svl_amd import pmlfile as P
    files = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "svl_runfiles", "*", "Partition",
                                          "*.json")))
    assert len(files) == 7
    for f in files:
        js = json.load(open(f))
        g = js["Global"]
        sim = js["Simulations"]["attributes"]
        if P.is_pml_run(f):
            r = P.read_pml_run(f)
            assert r.ntotal == g["ntotal"] and r.mesh.numbering()[3] == g["nfree"]
            assert not r.mesh.lumped or g.get("massform", "CONSISTENT").upper() == "LUMPED"
        else:
            r = S.read_run(f)
            assert r.model.nfree == g["nfree"] and r.ntotal == g["ntotal"]
            assert r.model.massform_lumped == (g.get("massform", "CONSISTENT").upper() == "LUMPED")
            assert r.algorithm == sim["algorithm"]["name"].upper()
            series = [ld["attributes"]["file"] for ld in js["Loads"].values() if "file" in ld["attributes"]]
            case = os.path.dirname(os.path.dirname(f))
            for s in set(series):
                n = int(float(open(os.path.join(case, s)).read().split()[0]))
                assert r.model.load_series.shape[1] == n
        assert r.integrator == sim["integrator"]["name"].upper() and r.dt == sim["integrator"]["dt"]
        assert r.nt == sim["analysis"]["nt"]
