"""SVL run files of PML models: PML2DQuad4 / PML3DHexa8 with lin2DQuad4 / lin3DHexa8 elements,
EQUAL constraints, concentrated point loads, ExtendedNewmarkBeta (or NewmarkBeta on 2D models,
whose PML element has no history matrix) with the Linear algorithm, NODE recorders.

The run-file layout is the one SVL's Driver reads (RP/12-Utilities/Driver.hpp):
``Nodes`` (ndof, freedof, totaldof, coords; a constrained unknown's freedof is the negative
constraint tag), ``Elements`` (PML2DQuad4 :1203-1220 with th, n, L, R, x0, npml; PML3DHexa8
:1288-1305 with n, L, R, x0, npml; lin2DQuad4 / lin3DHexa8 with material and th),
``Constraints`` (UpdateConstraints :440-505: the slave's total DOF, the masters' free DOFs and
the factors, as the pre-processor writes them, 01-Pre_Process/Core/SeismoVLAB.py:130-138).
The device supports one master per constrained unknown (EQUAL constraints).
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import numpy as np

from .pml import LIN2DQUAD4, LIN3DHEXA8, PML2DQUAD4, PML3DHEXA8, PmlMesh, PmlNewmark
from .svlfile import NodeRecorder, _fmt, _lc, _series_file

_CLASS = {"lin2dquad4": LIN2DQUAD4, "pml2dquad4": PML2DQUAD4, "lin3dhexa8": LIN3DHEXA8, "pml3dhexa8": PML3DHEXA8}
_NAME = {LIN2DQUAD4: "lin2DQuad4", PML2DQUAD4: "PML2DQuad4", LIN3DHEXA8: "lin3DHexa8", PML3DHEXA8: "PML3DHexa8"}


@dataclass
class PmlRun:
    mesh: PmlMesh
    node_tags: list                    # SVL tag of every node (mesh node order)
    totaldof: list                     # per node: its SVL total-DOF numbers
    ntotal: int
    integrator: str
    dt: float
    nt: int
    combo_name: str = "Combo"
    recorders: list = field(default_factory=list)
    skipped_recorders: list = field(default_factory=list)


def is_pml_run(path: str) -> bool:
    with open(path) as f:
        js = json.load(f)
    return any(_lc(e["name"]) in ("pml2dquad4", "pml3dhexa8") for e in js.get("Elements", {}).values())


def read_pml_run(path: str) -> PmlRun:
    base = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        js = json.load(f)
    g = js.get("Global", {})
    ndim = int(g.get("ndim", 0))
    if ndim not in (2, 3):
        raise ValueError("Global.ndim must be 2 or 3")
    for key in ("Supports", "Masses", "Sections"):
        if js.get(key):
            raise ValueError(f"run file has {key}: not on the device path")
    lumped = _lc(g.get("massform", "CONSISTENT")) == "lumped"
    tags = sorted(int(t) for t in js["Nodes"])
    nodes = [js["Nodes"][str(t)] for t in tags]
    index = {t: i for i, t in enumerate(tags)}
    # The device numbers total DOFs node after node in tag order; the file's must agree.
    ptr = np.concatenate([[0], np.cumsum([int(n["ndof"]) for n in nodes])]).astype(int)
    tot = [[int(v) for v in n["totaldof"]] for n in nodes]
    if [d for t in tot for d in t] != list(range(int(ptr[-1]))):
        raise ValueError("total DOFs must be numbered node after node in tag order")
    cons = {int(c): v for c, v in js.get("Constraints", {}).items()}
    fixed, equal_free = [], {}
    for i, n in enumerate(nodes):
        fd = [int(v) for v in n["freedof"]]
        fixed.append([v == -1 for v in fd])
        for a, v in enumerate(fd):
            if v < -1:
                c = cons.get(v)
                if c is None:
                    raise ValueError(f"node {tags[i]}: constraint {v} is not defined")
                if int(c["stag"]) != ptr[i] + a:
                    raise ValueError(f"constraint {v}: slave DOF {c['stag']} is not node {tags[i]} unknown {a}")
                if len(c["mtag"]) != 1:
                    raise ValueError(f"constraint {v}: one master per constrained unknown is on the device path")
                equal_free[(i, a)] = (int(c["mtag"][0]), float(c["factor"][0]))
    # free DOF -> (node, unknown) of the master
    owner = {}
    for i, n in enumerate(nodes):
        for a, v in enumerate(int(x) for x in n["freedof"]):
            if v >= 0:
                owner[v] = (i, a)
    if sorted(owner) != list(range(len(owner))):
        raise ValueError("free DOFs must be numbered 0..nfree-1")
    equal = [(s, a, owner[f][0], owner[f][1], fac) for (s, a), (f, fac) in sorted(equal_free.items())]
    mats = {int(t): v for t, v in js.get("Materials", {}).items()}
    cls, conn, E, nu, rho, th, pml = [], [], [], [], [], [], []
    for _, e in sorted(((int(t), v) for t, v in js.get("Elements", {}).items()), key=lambda x: x[0]):
        nm = _lc(e["name"])
        if nm not in _CLASS:
            raise ValueError(f"element {e['name']} is not on the PML device path")
        c = _CLASS[nm]
        if (ndim == 2) != (c in (LIN2DQUAD4, PML2DQUAD4)):
            raise ValueError(f"element {e['name']} does not match ndim {ndim}")
        at = e.get("attributes", {})
        nq = 4 if ndim == 2 else 8
        if _lc(at.get("rule", "GAUSS")) != "gauss" or int(at.get("np", nq)) != nq:
            raise ValueError(f"element {e['name']}: only the {nq}-point Gauss rule is on the PML device path")
        mt = mats[int(at["material"])]
        if _lc(mt["name"]) not in ("elastic2dplanestrain", "elastic3dlinear"):
            raise ValueError(f"material {mt['name']} is not on the PML device path")
        ma = mt["attributes"]
        cls.append(c)
        conn.append([index[int(t)] for t in e["conn"]])
        E.append(float(ma["E"]))
        nu.append(float(ma["nu"]))
        rho.append(float(ma.get("rho", 0.0)))
        th.append(float(at.get("th", 1.0)))
        if c in (PML2DQUAD4, PML3DHEXA8):
            x0 = [float(v) for v in at["x0"]] + [0.0] * (3 - ndim)
            npml = [float(v) for v in at["npml"]] + [0.0] * (3 - ndim)
            pml.append([float(at["n"]), float(at["L"]), float(at["R"])] + x0 + npml)
        else:
            pml.append([0.0] * 9)
    # The device PML path has no Rayleigh term (SVL adds am M_e + ak K_e to lin2DQuad4 /
    # lin3DHexa8 under Rayleigh damping, lin2DQuad4.cpp:352-364): refuse any that is not zero.
    for _, d in sorted(js.get("Dampings", {}).items(), key=lambda x: int(x[0])):
        dn, da = _lc(d["name"]), d.get("attributes", {})
        if dn == "free":
            continue
        if dn != "rayleigh":
            raise ValueError(f"damping {d['name']} is not on the PML device path")
        if (float(da.get("am", 0.0)) != 0.0 or float(da.get("ak", 0.0)) != 0.0) and da.get("list"):
            raise ValueError("Rayleigh damping is not on the PML device path")
    sim = js["Simulations"]
    sa = sim["attributes"]
    integ = _lc(sa["integrator"]["name"]).upper()
    if _lc(sa["algorithm"]["name"]) != "linear" or _lc(sa["analysis"]["name"]) != "dynamic":
        raise ValueError("PML models run with the Linear algorithm in a DYNAMIC analysis")
    has_g = PML3DHEXA8 in cls
    if not (integ == "EXTENDEDNEWMARK" or (integ == "NEWMARK" and not has_g)):
        raise ValueError(f"integrator {integ}: PML models need ExtendedNewmark (or Newmark without PML3DHexa8)")
    nt = int(sa["analysis"].get("nt", 1))
    combo = js.get("Combinations", {}).get(str(sim["combo"]), {})
    ca = combo.get("attributes", {})
    loads = []
    for lt, fac in zip(ca.get("load", []), ca.get("factor", [])):
        ld = js["Loads"][str(lt)]
        at = ld["attributes"]
        if _lc(ld["name"]) != "pointload" or _lc(at.get("type")) != "concentrated":
            raise ValueError(f"load {lt}: only CONCENTRATED point loads are on the PML device path")
        fn = _lc(at.get("name"))
        if fn == "timeseries":
            s = _series_file(at["file"], base)
        elif fn == "constant":
            s = np.full(nt, float(at["mag"]))
        else:
            raise ValueError(f"load {lt}: pattern {at.get('name')} is not on the device path")
        seen = set()
        for n in at["list"]:
            if int(n) in seen:
                continue
            seen.add(int(n))
            for c, dv in enumerate(float(v) for v in at["dir"][:ndim]):
                if dv != 0.0:
                    loads.append((index[int(n)], c, float(fac) * dv, s))
    mesh = PmlMesh(np.array([[float(v) for v in n["coords"][:ndim]] for n in nodes]),
                   np.array([int(n["ndof"]) for n in nodes]), fixed, np.array(cls), np.array(conn), np.array(E),
                   np.array(nu), np.array(rho), np.array(th), np.array(pml), equal=equal, loads=loads, lumped=lumped)
    run = PmlRun(mesh=mesh, node_tags=tags, totaldof=tot, ntotal=int(ptr[-1]), integrator=integ,
                 dt=float(sa["integrator"]["dt"]), nt=nt, combo_name=combo.get("name", "Combo"))
    for _, r in sorted(js.get("Recorders", {}).items(), key=lambda x: int(x[0])):
        if _lc(r["name"]) == "node" and _lc(r.get("resp", "")) in ("disp", "vel", "accel"):
            run.recorders.append(NodeRecorder(r["file"], r["resp"].upper(), [int(t) for t in r["list"]],
                                              int(r.get("ndps", 8))))
        else:
            run.skipped_recorders.append(r)
    return run


def write_pml_run(mesh: PmlMesh, path: str, dt: float, nt: int, integrator: str = "EXTENDEDNEWMARK",
                  recorders=()) -> list:
    """Writes `mesh` as an SVL run file (node i has tag i + 1, element e tag e + 1, constraint j
    tag -(j + 2), one material per element, one TIMESERIES point load per mesh load with its
    series in a text file next to the run file). recorders: (file, resp, [node tags], ndps)."""
    nd = mesh.ndim
    ptr, free, fac, nfree = mesh.numbering()
    nodes, cons = {}, {}
    ctag = {}
    for j, (s, a, m, b, f) in enumerate(mesh.equal):
        ctag[(s, a)] = -(j + 2)
        cons[str(-(j + 2))] = {"stag": int(ptr[s] + a), "mtag": [int(free[ptr[m] + b])], "factor": [float(f)]}
    for i in range(len(mesh.ndof)):
        fd = []
        for a in range(int(mesh.ndof[i])):
            fd.append(ctag[(i, a)] if (i, a) in ctag else int(free[ptr[i] + a]))
        nodes[str(i + 1)] = {"ndof": int(mesh.ndof[i]), "freedof": fd,
                             "totaldof": list(range(int(ptr[i]), int(ptr[i + 1]))),
                             "coords": [float(v) for v in mesh.coords[i]]}
    mats, elems = {}, {}
    for e in range(len(mesh.conn)):
        mats[str(e + 1)] = {"name": "Elastic2DPlaneStrain" if nd == 2 else "Elastic3DLinear",
                            "attributes": {"E": float(mesh.E[e]), "nu": float(mesh.nu[e]), "rho": float(mesh.rho[e])}}
        c = int(mesh.elem_class[e])
        at = {"material": e + 1, "rule": "GAUSS", "np": 4 if nd == 2 else 8}
        if nd == 2:
            at["th"] = float(mesh.th[e])
        if c in (PML2DQUAD4, PML3DHEXA8):
            q = [float(v) for v in mesh.pml[e]]
            at.update({"n": q[0], "L": q[1], "R": q[2], "x0": q[3:3 + nd], "npml": q[6:6 + nd]})
        elems[str(e + 1)] = {"name": _NAME[c].upper(), "conn": [int(n) + 1 for n in mesh.conn[e]], "attributes": at}
    base = os.path.dirname(os.path.abspath(path))
    stem = os.path.splitext(os.path.basename(path))[0]
    loads, lids, facs = {}, [], []
    for l, (n, a, f, s) in enumerate(mesh.loads):
        fname = os.path.join(base, f"{stem}_load{l + 1}.txt")
        s = np.asarray(s, np.float64)
        with open(fname, "w") as fh:
            fh.write(f"{s.size}\n" + "\n".join(repr(float(v)) for v in s) + "\n")
        dvec = [0.0] * nd
        dvec[a] = 1.0
        loads[str(l + 1)] = {"name": "PointLoad", "attributes": {"type": "CONCENTRATED", "name": "TIMESERIES",
                                                                 "file": fname, "list": [int(n) + 1], "dir": dvec}}
        lids.append(l + 1)
        facs.append(float(f))
    out = {"Global": {"ndim": nd, "massform": "LUMPED" if mesh.lumped else "CONSISTENT", "update": "RESTART",
                      "ntotal": int(ptr[-1]), "nfree": int(nfree)},
           "Nodes": nodes, "Materials": mats, "Elements": elems, "Constraints": cons, "Loads": loads,
           "Combinations": {"1": {"name": "Combo", "attributes": {"load": lids, "factor": facs}}},
           "Simulations": {"combo": 1, "attributes": {
               "solver": {"name": "EIGEN", "update": "OFF"},
               "algorithm": {"name": "LINEAR", "nstep": 1},
               "integrator": {"name": integrator, "dt": dt},
               "analysis": {"name": "DYNAMIC", "nt": nt}}},
           "Recorders": {str(i + 1): {"name": "NODE", "file": f, "resp": r, "ndps": int(p), "nsamp": 1,
                                      "list": [int(n) for n in lst]}
                         for i, (f, r, lst, p) in enumerate(recorders)}}
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    return list(range(1, len(mesh.ndof) + 1))


def write_node_recorder_mixed(path: str, rec: NodeRecorder, tags: list, totaldof: list, ntotal: int, nt: int,
                              values: np.ndarray):
    """NODE recorder file for nodes of different unknown counts: values[k-1, j] is the j-th
    recorded component after step k, the recorder's nodes one after the other."""
    rows = [tags.index(t) for t in rec.tags]
    ncomp = sum(len(totaldof[r]) for r in rows)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write(f"{len(rec.tags)} {ncomp} {ntotal} {nt}\n")
        for t, r in zip(rec.tags, rows):
            f.write(f"{t} {len(totaldof[r])}" + "".join(f" {int(d)}" for d in totaldof[r]) + "\n")
        for k in range(values.shape[0]):
            f.write("".join(_fmt(float(v), rec.ndps) + " " for v in values[k]) + "\n")


def run_pml(pr: PmlRun, out_dir: str | None = None, json_path: str | None = None, rtol: float = 1e-14) -> dict:
    """Runs a PML run file on the device (steps k = 1..nt-1) and writes its NODE recorders
    (default: SVL's <run dir>/../Solution/<combo>/). Returns {file: values [nt-1, components]}."""
    tags = pr.node_tags
    for r in pr.recorders:
        missing = [t for t in r.tags if t not in tags]
        if missing:
            raise ValueError(f"recorder nodes {missing[:5]} are not in the model")
    needed = sorted({t for r in pr.recorders for t in r.tags})
    rec = [(tags.index(t), a) for t in needed for a in range(int(pr.mesh.ndof[tags.index(t)]))]
    col = {pair: j for j, pair in enumerate(rec)}
    pm = PmlNewmark(pr.mesh, pr.dt, rtol=rtol)
    hu, hv, ha, _ = pm.advance(1, pr.nt - 1, rec)
    resp = {"DISP": hu, "VEL": hv, "ACCEL": ha}
    base = out_dir
    if base is None:
        root = os.path.dirname(os.path.abspath(json_path)) if json_path else "."
        base = os.path.join(root, "..", "Solution", pr.combo_name)
    out = {}
    for r in pr.recorders:
        cols = [col[(tags.index(t), a)] for t in r.tags for a in range(int(pr.mesh.ndof[tags.index(t)]))]
        vals = resp[r.response][:, cols]
        p = os.path.join(base, r.file)
        write_node_recorder_mixed(p, r, tags, pr.totaldof, pr.ntotal, pr.nt, vals)
        out[p] = vals
    return out
