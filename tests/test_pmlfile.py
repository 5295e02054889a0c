"""CPU checks of PML run files: write -> read reproduces the mesh, SVL's numbering with its
negative constraint tags, the EQUAL list and the loads; refused contents raise."""
import json

import numpy as np
import pytest

from svl_amd.pml import halfspace_box3d, soil_column, soil_rod3d
from svl_amd.pmlfile import is_pml_run, read_pml_run, write_pml_run


@pytest.mark.parametrize("build", [lambda s: soil_column(20, 5, s), lambda s: soil_rod3d(12, 4, s),
                                   lambda s: halfspace_box3d(2, 2, 2, 1, load_series=s)])
def test_round_trip(tmp_path, build):
    s = np.sin(np.arange(30) * 0.3)
    m = build(s)
    path = str(tmp_path / "run.json")
    write_pml_run(m, path, 0.01, 30, recorders=[("disp.out", "DISP", [1, len(m.ndof)], 8)])
    assert is_pml_run(path)
    r = read_pml_run(path)
    q = r.mesh
    for name in ("coords", "ndof", "elem_class", "conn", "E", "nu", "rho"):
        np.testing.assert_array_equal(getattr(q, name), getattr(m, name), err_msg=name)
    is_pml = (m.elem_class == 1) | (m.elem_class == 3)      # regular elements carry no PML parameters
    np.testing.assert_array_equal(q.pml[is_pml], m.pml[is_pml])
    assert q.fixed == [list(f) for f in m.fixed]
    assert sorted(q.equal) == sorted(m.equal)
    for a, b in zip(m.numbering(), q.numbering()):
        np.testing.assert_array_equal(a, b)
    assert [(n, c, f) for n, c, f, _ in q.loads] == [(n, c, f) for n, c, f, _ in m.loads]
    for (_, _, _, s1), (_, _, _, s2) in zip(q.loads, m.loads):
        np.testing.assert_array_equal(s1, s2)
    assert r.nt == 30 and r.dt == 0.01 and r.recorders[0].tags == [1, len(m.ndof)]
    js = json.load(open(path))
    assert all(int(v) < -1 for n in js["Nodes"].values() for v in n["freedof"] if int(v) < -1)


def test_refuses_newmark_with_history_matrix(tmp_path):
    m = soil_rod3d(4, 2, np.ones(5))
    path = str(tmp_path / "run.json")
    write_pml_run(m, path, 0.01, 5, integrator="NEWMARK")
    with pytest.raises(ValueError, match="ExtendedNewmark"):
        read_pml_run(path)


def test_refuses_multi_master_constraint(tmp_path):
    m = soil_column(4, 2, np.ones(5))
    path = str(tmp_path / "run.json")
    write_pml_run(m, path, 0.01, 5)
    js = json.load(open(path))
    c = next(iter(js["Constraints"].values()))
    c["mtag"].append(0)
    c["factor"].append(0.5)
    json.dump(js, open(path, "w"))
    with pytest.raises(ValueError, match="one master"):
        read_pml_run(path)


def test_refuses_rayleigh_damping(tmp_path):
    """The device PML path has no Rayleigh term: a run file whose Dampings give the elements a
    nonzero am or ak is refused; Free damping (or Rayleigh with zero coefficients) passes."""
    m = soil_column(4, 2, np.ones(5))
    path = str(tmp_path / "run.json")
    write_pml_run(m, path, 0.01, 5)
    js = json.load(open(path))
    tags = [int(t) for t in js["Elements"]]
    js["Dampings"] = {"1": {"name": "Rayleigh", "attributes": {"am": 0.0, "ak": 0.0, "list": tags}}}
    json.dump(js, open(path, "w"))
    read_pml_run(path)
    js["Dampings"]["1"]["attributes"]["ak"] = 1e-3
    json.dump(js, open(path, "w"))
    with pytest.raises(ValueError, match="Rayleigh"):
        read_pml_run(path)


@pytest.mark.parametrize("attr", [{"rule": "LOBATTO"}, {"np": 9}])
def test_refuses_other_quadrature(tmp_path, attr):
    m = soil_column(4, 2, np.ones(5))
    path = str(tmp_path / "run.json")
    write_pml_run(m, path, 0.01, 5)
    js = json.load(open(path))
    e = next(iter(js["Elements"].values()))
    e.setdefault("attributes", {}).update(attr)
    json.dump(js, open(path, "w"))
    with pytest.raises(ValueError, match="Gauss rule"):
        read_pml_run(path)
