"""SVL run files (JSON) in, SVL NODE recorder files out, for the models the device path runs.

Reader (``read_run``) of the run-file schema that SVL's driver parses
(RP/12-Utilities/Driver.hpp): ``Global`` (ndim, massform :1993-2004), ``Nodes`` (ndof, freedof,
totaldof, coords :304-345), ``Materials`` (Elastic2DPlaneStrain, Elastic3DLinear,
PlasticPlaneStrainJ2 / Plastic3DJ2 with K, G, rho, h, Sy, beta, Viscous1DLinear :567-760),
``Elements`` (lin2DQuad4 :1185-1193, lin3DHexa8 :1272-1279, ZeroLength1D :1072-1078),
``Dampings`` (Free / Rayleigh with an element list :1362-1454), ``Loads`` (PointLoad,
CONCENTRATED, CONSTANT / TIMESERIES with a node list and a series file whose first entry is
the number of samples :1461-1560), ``Combinations`` (:1929-1975), ``Simulations`` (solver,
algorithm, integrator, analysis, combo :1747-1850) and ``Recorders`` (:1858-1927).
Continuum nodes become mesh nodes in ascending tag order; a ZeroLength1D element with a
Viscous1DLinear material between a mesh node and a fully restrained node is a Lysmer dashpot;
the free-DOF numbering is taken from the file. Anything else (other elements or materials,
constraints, point masses, body loads) is refused with ValueError: it is not on the device
path. ``ElementLoad`` GENERALWAVE (domain reduction, :1669-1720) is read with its per-node
signal files; ``Supports`` (CONSTANT value or TIMESERIES file per DOF, :505-563) with the
combination's ``SupportMotion`` loads (:1725-1734) become the model's support motion (restrained
DOFs of mesh nodes, Linear algorithm).

Writer (``write_run``) of the same schema from a ``Model`` (the inverse, with a restrained
copy node per dashpot as setLysmerDomain makes them, PP/Method/Builder.py:1086-1131), so a
model built here can be handed to SVL's own driver.

NODE recorder (``write_node_recorder``): SVL's text format (RP/12-Utilities/Recorder.cpp:
73-106 header, 239-268 rows): ``nIDs nDofs ntotal nsteps``, then per node ``tag ndof totaldof...``,
then one row per step k = 1..nt-1 with every DOF of every listed node in scientific notation
with ``ndps`` digits, each value followed by a space.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import numpy as np

from .mesh import ELEM_HEXA8, ELEM_QUAD4, MAT_J2, Model

_CONTINUUM = {"lin2dquad4": (ELEM_QUAD4, 2, 4), "lin3dhexa8": (ELEM_HEXA8, 3, 8)}
_ELASTIC = {"elastic2dplanestrain": 2, "elastic3dlinear": 3}
_J2 = {"plasticplanestrainj2": 2, "plastic3dj2": 3}


@dataclass
class NodeRecorder:
    file: str
    response: str          # DISP, VEL, ACCEL
    tags: list             # SVL node tags
    ndps: int


@dataclass
class SvlRun:
    model: Model
    node_tags: np.ndarray            # SVL tag of every mesh node (model node order)
    totaldof: np.ndarray             # [nnodes, ndim] SVL total-DOF numbers of the mesh nodes
    ntotal: int
    integrator: str                  # CENTRALDIFFERENCE, EXTENDEDNEWMARK, NEWMARK
    algorithm: str                   # LINEAR, NEWTON
    dt: float
    nt: int
    cnvgtol: float = 1e-6
    nmax: int = 50
    cnvgtest: int = 4
    combo_name: str = "Combo"
    recorders: list = field(default_factory=list)
    skipped_recorders: list = field(default_factory=list)
    # Restrained nodes outside the mesh (dashpot partners): tag -> (totaldof, [(mesh node, dir, eta)]).
    partners: dict = field(default_factory=dict)


def _lc(s):
    return str(s).lower()


def _resolve(path, base):
    """A data file named in a run file. SVL opens it as written, relative to the directory the
    run starts in: the case directory, one level above the Partition/ folder that SVL's
    pre-processor writes the run files to. Relative names are looked up next to the run file,
    then in that parent folder."""
    if os.path.isabs(path):
        return path
    for d in (base, os.path.dirname(base)):
        p = os.path.join(d, path)
        if os.path.exists(p):
            return p
    return os.path.join(base, path)


def _series_file(path, base):
    p = _resolve(path, base)
    with open(p) as f:
        vals = f.read().split()
    n = int(float(vals[0]))
    return np.asarray([float(v) for v in vals[1:1 + n]], dtype=np.float64)


def _read_drm(model: Model, at: dict, etag_index: dict, index: dict, base: str, factor: float):
    """ElementLoad GENERALWAVE (Driver.hpp:1669-1720): the listed elements, and per node of
    those elements (ascending tag) a file named by the load's pattern with '$' replaced by the
    node tag, holding ``nt nFields cond`` and nt rows of U, V, A components."""
    elems = []
    for t in at["list"]:
        if int(t) not in etag_index:
            raise ValueError(f"DRM element {t} is not a continuum element")
        elems.append(etag_index[int(t)])
    elems = np.asarray(elems, np.int32)
    tags = sorted({int(model.node_tags[n]) for n in model.conn[elems].reshape(-1)})
    nf = 3 * model.ndim
    sig, cond = [], []
    for t in tags:
        fname = str(at["file"]).replace("$", str(t), 1)
        p = _resolve(fname, base)
        with open(p) as f:
            vals = f.read().split()
        nt, nfields, c = int(vals[0]), int(vals[1]), int(vals[2])
        if nfields != nf:
            raise ValueError(f"DRM node {t}: {nfields} fields, the device path reads U, V, A ({nf})")
        sig.append(np.asarray([float(v) for v in vals[3:3 + nt * nf]], np.float64).reshape(nt, nf))
        cond.append(1 if c else 0)
    nt = max(s.shape[0] for s in sig)
    signal = np.zeros((len(tags), nt, nf))
    for i, s in enumerate(sig):
        signal[i, :s.shape[0]] = s
    from .drm import attach
    attach(model, elems, np.asarray([index[t] for t in tags], np.int32), np.asarray(cond, np.int32), signal, factor)


def _read_support_motion(model: Model, at: dict, supports: dict, index: dict, base: str, factor: float):
    """A SupportMotion load (Driver.hpp:1725-1734): every listed node's Supports entry
    (Driver.hpp:505-563: per listed DOF a CONSTANT value or a TIMESERIES file holding the number
    of samples and the samples) with the load's combination factor. A node without a Supports
    entry moves nothing (Node::GetSupportMotion is zero there)."""
    for t in at.get("list", []):
        t = int(t)
        if t not in supports:
            continue
        if t not in index:
            raise ValueError(f"support motion at node {t}: not a node of the continuum mesh")
        sp = supports[t]
        kind = _lc(sp.get("type", ""))
        per_dof = {}
        for i, dof in enumerate(sp.get("dof", [])):
            if kind == "constant":
                per_dof[int(dof)] = np.asarray([float(sp["value"][i])])
            elif kind == "timeseries":
                per_dof[int(dof)] = _series_file(sp["file"][i], base)
            else:
                raise ValueError(f"support motion at node {t}: type {sp.get('type')} is not CONSTANT or TIMESERIES")
        for dof, g in sorted(per_dof.items()):  # Node::SupportMotion is a map ordered by DOF
            if dof < 0 or dof >= model.ndim:
                raise ValueError(f"support motion at node {t}: DOF {dof} out of range")
            if not model.restrained[index[t], dof]:
                raise ValueError(f"support motion at node {t}, DOF {dof}: a free DOF is not on the device path")
            model.add_support_motion(index[t], dof, g, factor)


def read_run(path: str) -> SvlRun:
    """Reads an SVL run file into a device-path model (see the module docstring)."""
    base = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        js = json.load(f)
    g = js.get("Global", {})
    ndim = int(g.get("ndim", 0))
    lumped = _lc(g.get("massform", "CONSISTENT")) == "lumped"
    for key in ("Constraints", "Masses", "Sections"):
        if js.get(key):
            raise ValueError(f"run file has {key}: not on the device path")
    mats = {int(t): v for t, v in js.get("Materials", {}).items()}
    nodes = {int(t): v for t, v in js.get("Nodes", {}).items()}
    elems = sorted(((int(t), v) for t, v in js.get("Elements", {}).items()), key=lambda x: x[0])
    cont, dash = [], []
    etype = None
    for tag, e in elems:
        name = _lc(e["name"])
        at = e.get("attributes", {})
        if name in _CONTINUUM:
            t, d, nen = _CONTINUUM[name]
            if etype is None:
                etype = t
            elif etype != t:
                raise ValueError("run file mixes continuum element types")
            if d != ndim or len(e["conn"]) != nen:
                raise ValueError(f"element {tag}: {e['name']} does not match ndim {ndim}")
            if _lc(at.get("rule", "GAUSS")) != "gauss" or int(at.get("np", nen)) != nen:
                raise ValueError(f"element {tag}: only the {nen}-point Gauss rule is on the device path")
            cont.append((tag, e))
        elif name == "zerolength1d":
            dash.append((tag, e))
        else:
            raise ValueError(f"element {tag}: {e['name']} is not on the device path")
    if not cont:
        raise ValueError("run file has no lin2DQuad4 / lin3DHexa8 elements")
    mesh_tags = sorted({int(n) for _, e in cont for n in e["conn"]})
    index = {t: i for i, t in enumerate(mesh_tags)}
    coords = np.array([[float(c) for c in nodes[t]["coords"][:ndim]] for t in mesh_tags])
    freedof = np.array([[int(v) for v in nodes[t]["freedof"][:ndim]] for t in mesh_tags], dtype=np.int32)
    totaldof = np.array([[int(v) for v in nodes[t]["totaldof"][:ndim]] for t in mesh_tags], dtype=np.int32)
    conn = np.array([[index[int(n)] for n in e["conn"]] for _, e in cont], dtype=np.int32)
    ne = len(cont)
    E, nu, rho = np.zeros(ne), np.zeros(ne), np.zeros(ne)
    matt = np.zeros(ne, np.int32)
    j2 = np.zeros((ne, 5))
    th = 1.0
    for i, (tag, e) in enumerate(cont):
        at = e["attributes"]
        if etype == ELEM_QUAD4:
            th = float(at.get("th", 1.0)) if i == 0 else th
            if float(at.get("th", 1.0)) != th:
                raise ValueError("lin2DQuad4 elements with different thicknesses")
        m = mats[int(at["material"])]
        mname, ma = _lc(m["name"]), m.get("attributes", {})
        if mname in _ELASTIC and _ELASTIC[mname] == ndim:
            E[i], nu[i], rho[i] = float(ma.get("E", 0.0)), float(ma.get("nu", 0.0)), float(ma.get("rho", 0.0))
        elif mname in _J2 and _J2[mname] == ndim:
            K, G = float(ma.get("K", 0.0)), float(ma.get("G", 0.0))
            matt[i] = MAT_J2
            j2[i] = [K, G, float(ma.get("Sy", 0.0)), float(ma.get("h", 0.0)), float(ma.get("beta", 0.0))]
            E[i] = 9.0 * K * G / (3.0 * K + G)
            nu[i] = (3.0 * K - 2.0 * G) / (2.0 * (3.0 * K + G))
            rho[i] = float(ma.get("rho", 0.0))
        else:
            raise ValueError(f"element {tag}: material {m['name']} is not on the device path")
    dn, dd, de = [], [], []
    partner_of = {}
    for tag, e in dash:
        a, b = (int(n) for n in e["conn"])
        at = e["attributes"]
        m = mats[int(at["material"])]
        if _lc(m["name"]) != "viscous1dlinear":
            raise ValueError(f"ZeroLength1D {tag}: only Viscous1DLinear dashpots are on the device path")
        pairs = [(a, b), (b, a)]
        mesh_end = [p for p in pairs if p[0] in index and all(int(v) == -1 for v in nodes[p[1]]["freedof"])]
        if not mesh_end:
            raise ValueError(f"ZeroLength1D {tag}: a dashpot must join a mesh node to a fully restrained node")
        dn.append(index[mesh_end[0][0]])
        dd.append(int(at["dir"]))
        de.append(float(m["attributes"]["eta"]))
        pt = mesh_end[0][1]
        partner_of.setdefault(pt, ([int(v) for v in nodes[pt]["totaldof"][:ndim]], []))[1].append(
            (dn[-1], dd[-1], de[-1]))
    model = Model(ndim=ndim, coords=coords, node_tags=np.asarray(mesh_tags, dtype=np.int64), elem_type=etype,
                  conn=conn, E=E, nu=nu, rho=rho, thickness=th, restrained=freedof < 0, mat_type=matt, j2=j2,
                  dash_node=np.asarray(dn, np.int32), dash_dir=np.asarray(dd, np.int32),
                  dash_eta=np.asarray(de, np.float64), massform_lumped=lumped,
                  name=os.path.splitext(os.path.basename(path))[0])
    model.free_dof_override = freedof
    nfree = int(freedof.max()) + 1 if (freedof >= 0).any() else 0
    if sorted(freedof[freedof >= 0].tolist()) != list(range(nfree)):
        raise ValueError("the free DOFs of the mesh nodes must be numbered 0..nfree-1")
    # Dampings (per continuum element; dashpot elements have no Rayleigh term).
    am, ak = np.zeros(ne), np.zeros(ne)
    etag_index = {t: i for i, (t, _) in enumerate(cont)}
    for _, d in sorted(js.get("Dampings", {}).items(), key=lambda x: int(x[0])):
        nm, at = _lc(d["name"]), d.get("attributes", {})
        if nm not in ("free", "rayleigh"):
            raise ValueError(f"damping {d['name']} is not on the device path")
        for t in at.get("list", []):
            if int(t) in etag_index:
                i = etag_index[int(t)]
                am[i] = float(at.get("am", 0.0)) if nm == "rayleigh" else 0.0
                ak[i] = float(at.get("ak", 0.0)) if nm == "rayleigh" else 0.0
    model.ray_am, model.ray_ak = am, ak
    # Simulation.
    sim = js["Simulations"]
    sa = sim["attributes"]
    integ = _lc(sa["integrator"]["name"]).upper()
    alg = sa["algorithm"]
    run = SvlRun(model=model, node_tags=np.asarray(mesh_tags), totaldof=totaldof,
                 ntotal=int(g.get("ntotal", 0)), integrator=integ, algorithm=_lc(alg["name"]).upper(),
                 dt=float(sa["integrator"].get("dt", 0.0)), nt=int(sa["analysis"].get("nt", 1)),
                 cnvgtol=float(alg.get("cnvgtol", 1e-6)), nmax=int(alg.get("nstep", 50)),
                 cnvgtest=int(alg.get("cnvgtest", 4)))
    run.partners = partner_of
    if _lc(sa["analysis"]["name"]) != "dynamic":
        raise ValueError("only DYNAMIC analyses are on the device path")
    # Loads of the simulation's combination.
    combo = js.get("Combinations", {}).get(str(sim["combo"]), {})
    run.combo_name = combo.get("name", "Combo")
    ca = combo.get("attributes", {})
    loads = js.get("Loads", {})
    series, rows = [], []
    supports = {int(t): v for t, v in js.get("Supports", {}).items()}
    for lt, fac in zip(ca.get("load", []), ca.get("factor", [])):
        ld = loads[str(lt)]
        at = ld["attributes"]
        if _lc(ld["name"]) == "supportmotion":
            _read_support_motion(model, at, supports, index, base, float(fac))
            continue
        if _lc(ld["name"]) == "elementload" and _lc(at.get("type")) == "generalwave":
            if model.drm_elem is not None:
                raise ValueError(f"load {lt}: one domain reduction load per combination is on the device path")
            _read_drm(model, at, etag_index, index, base, float(fac))
            continue
        if _lc(ld["name"]) != "pointload":
            raise ValueError(f"load {lt}: {ld['name']} is not on the device path")
        if _lc(at.get("type")) != "concentrated":
            raise ValueError(f"load {lt}: only CONCENTRATED point loads are on the device path")
        fn = _lc(at.get("name"))
        if fn == "timeseries":
            s = _series_file(at["file"], base)
        elif fn == "constant":
            s = np.full(run.nt, float(at["mag"]))
        else:
            raise ValueError(f"load {lt}: pattern {at.get('name')} is not on the device path")
        direction = [float(v) for v in at["dir"]][:ndim]
        seen = set()
        for n in at["list"]:
            if int(n) in seen:  # SVL assigns per node: a repeated node counts once
                continue
            seen.add(int(n))
            rows.append((index[int(n)], direction, float(fac)))
            series.append(s)
    if rows:
        nts = max(len(s) for s in series)
        model.load_node = np.asarray([r[0] for r in rows], np.int32)
        model.load_dir = np.asarray([r[1] for r in rows], np.float64)
        model.load_factor = np.asarray([r[2] for r in rows], np.float64)
        model.load_series = np.zeros((len(rows), nts))
        for i, s in enumerate(series):
            model.load_series[i, :len(s)] = s
    if len(model.sm_node):
        if run.algorithm != "LINEAR":
            raise ValueError("support motion with the Newton-Raphson algorithm is not on the device path")
        # NewmarkBeta applies the increment of sample k - 1 at step k (NewmarkBeta.cpp:187).
        model.sm_offset = 1 if integ == "NEWMARK" else 0
    # NODE recorders (others are listed as skipped).
    for _, r in sorted(js.get("Recorders", {}).items(), key=lambda x: int(x[0])):
        if _lc(r["name"]) == "node" and _lc(r.get("resp", "")) in ("disp", "vel", "accel", "reaction"):
            run.recorders.append(NodeRecorder(r["file"], r["resp"].upper(), [int(t) for t in r["list"]],
                                              int(r.get("ndps", 8))))
        else:
            run.skipped_recorders.append(r)
    return run


def write_run(model: Model, path: str, integrator: str, dt: float, nt: int, algorithm: str = "LINEAR",
              cnvgtol: float = 1e-6, nmax: int = 50, cnvgtest: int = 4, recorders=(), series_dir: str | None = None):
    """Writes `model` as an SVL run file (tags: mesh nodes 1..n, dashpot partner nodes after
    them; continuum elements 1..ne, dashpots after them; one material per distinct parameter
    set; one TIMESERIES PointLoad per model load with its series in a text file next to the
    run file). recorders: (file, resp, [node tags], ndps); mesh node i has tag i + 1, the
    restrained copy of the j-th dashpot node (order of first appearance) tag nnodes + j + 1.
    Returns the mesh node tags."""
    nd = model.ndim
    out = {"Global": {"ndim": nd, "massform": "LUMPED" if model.massform_lumped else "CONSISTENT",
                      "update": "RESTART"}}
    nn, ne = model.nnodes, model.nelem
    fd = model.free_dof()
    nodes = {}
    for i in range(nn):
        tot = [i * nd + c for c in range(nd)]
        nodes[str(i + 1)] = {"ndof": nd, "freedof": [int(v) for v in fd[i]], "totaldof": tot,
                             "coords": [float(c) for c in model.coords[i]]}
    mats, mat_of = {}, {}

    def material(key, name, attrs):
        if key not in mat_of:
            mat_of[key] = len(mats) + 1
            mats[str(mat_of[key])] = {"name": name, "attributes": attrs}
        return mat_of[key]

    elems = {}
    ename = "lin2DQuad4" if model.elem_type == ELEM_QUAD4 else "lin3DHexa8"
    for e in range(ne):
        if model.mat_type[e] == MAT_J2:
            K, G, sy, H, beta = (float(v) for v in model.j2[e])
            mname = "PlasticPlaneStrainJ2" if nd == 2 else "Plastic3DJ2"
            mt = material(("j2", K, G, sy, H, beta, float(model.rho[e])), mname,
                          {"K": K, "G": G, "rho": float(model.rho[e]), "h": H, "Sy": sy, "beta": beta})
        else:
            mname = "Elastic2DPlaneStrain" if nd == 2 else "Elastic3DLinear"
            mt = material(("el", float(model.E[e]), float(model.nu[e]), float(model.rho[e])), mname,
                          {"E": float(model.E[e]), "nu": float(model.nu[e]), "rho": float(model.rho[e])})
        at = {"material": mt, "rule": "GAUSS", "np": 4 if nd == 2 else 8}
        if nd == 2:
            at["th"] = float(model.thickness)
        elems[str(e + 1)] = {"name": ename, "conn": [int(c) + 1 for c in model.conn[e]], "attributes": at}
    ntot = nn * nd
    partner = {}
    for k in range(model.dash_node.size):
        n = int(model.dash_node[k])
        if n not in partner:  # one restrained copy per dashpot node (setLysmerDomain)
            partner[n] = nn + len(partner) + 1
            tag = partner[n]
            nodes[str(tag)] = {"ndof": nd, "freedof": [-1] * nd,
                               "totaldof": [ntot + (tag - nn - 1) * nd + c for c in range(nd)],
                               "coords": [float(c) for c in model.coords[n]]}
        mt = material(("visc", float(model.dash_eta[k])), "Viscous1DLinear", {"eta": float(model.dash_eta[k])})
        elems[str(ne + k + 1)] = {"name": "ZeroLength1D", "conn": [n + 1, partner[n]],
                                  "attributes": {"material": mt, "dir": int(model.dash_dir[k])}}
    out["Global"]["ntotal"] = ntot + len(partner) * nd
    out["Global"]["nfree"] = model.nfree
    out["Nodes"], out["Materials"], out["Elements"] = nodes, mats, elems
    damp = {}
    am = np.zeros(ne) if model.ray_am is None else np.asarray(model.ray_am)
    ak = np.zeros(ne) if model.ray_ak is None else np.asarray(model.ray_ak)
    groups = {}
    for e in range(ne):
        groups.setdefault((float(am[e]), float(ak[e])), []).append(e + 1)
    for i, ((a, b), lst) in enumerate(sorted(groups.items())):
        damp[str(i + 1)] = ({"name": "Free", "attributes": {"list": lst}} if a == 0.0 and b == 0.0 else
                            {"name": "Rayleigh", "attributes": {"am": a, "ak": b, "list": lst}})
    out["Dampings"] = damp
    base = series_dir or os.path.dirname(os.path.abspath(path))
    os.makedirs(base, exist_ok=True)
    loads, lids, facs = {}, [], []
    stem = os.path.splitext(os.path.basename(path))[0]
    for l in range(model.load_node.size):
        fname = os.path.join(base, f"{stem}_load{l + 1}.txt")
        s = model.load_series[l]
        with open(fname, "w") as f:
            f.write(f"{s.size}\n")
            f.write("\n".join(repr(float(v)) for v in s))
            f.write("\n")
        loads[str(l + 1)] = {"name": "PointLoad", "attributes": {
            "type": "CONCENTRATED", "name": "TIMESERIES", "file": fname, "list": [int(model.load_node[l]) + 1],
            "dir": [float(v) for v in model.load_dir[l]]}}
        lids.append(l + 1)
        facs.append(float(model.load_factor[l]))
    if model.drm_elem is not None and len(model.drm_elem):
        # ElementLoad GENERALWAVE with one signal file per DRM node ('$' = node tag).
        pattern = os.path.join(base, f"{stem}_drm_$.txt")
        for i, n in enumerate(model.drm_node):
            sg = model.drm_signal[i]
            with open(pattern.replace("$", str(int(n) + 1)), "w") as f:
                f.write(f"{sg.shape[0]} {sg.shape[1]} {int(model.drm_cond[i])}\n")
                for row in sg:
                    f.write(" ".join(repr(float(v)) for v in row) + "\n")
        lid = len(loads) + 1
        loads[str(lid)] = {"name": "ElementLoad", "attributes": {
            "type": "GENERALWAVE", "name": "TIMESERIES", "file": pattern,
            "list": [int(e) + 1 for e in model.drm_elem]}}
        lids.append(lid)
        facs.append(float(model.drm_factor))
    if len(model.sm_node):
        # Supports (one TIMESERIES file per node DOF) and one SupportMotion load per distinct
        # factor listing its nodes; a load moves every DOF of its nodes, so each node's DOFs must
        # carry the same factors and each (node, DOF) one series.
        per_node = {}
        for j in range(len(model.sm_node)):
            n, c, f = int(model.sm_node[j]), int(model.sm_dir[j]), float(model.sm_factor[j])
            dofs = per_node.setdefault(n, {})
            g = np.asarray(model.sm_series[j], np.float64)
            if c in dofs and not np.array_equal(dofs[c][0], g):
                raise ValueError(f"node {n}, DOF {c}: SVL holds one support series per DOF")
            dofs.setdefault(c, (g, []))[1].append(f)
        sup, by_factor = {}, {}
        for n, dofs in sorted(per_node.items()):
            fsets = {tuple(v[1]) for v in dofs.values()}
            if len(fsets) != 1:
                raise ValueError(f"node {n}: its DOFs carry different support-motion factors")
            files = []
            for c, (g, _) in sorted(dofs.items()):
                fname = os.path.join(base, f"{stem}_support{n + 1}_{c}.txt")
                with open(fname, "w") as f:
                    f.write(f"{g.size}\n")
                    f.write("\n".join(repr(float(v)) for v in g))
                    f.write("\n")
                files.append(fname)
            sup[str(n + 1)] = {"type": "TIMESERIES", "file": files, "dof": sorted(dofs)}
            for f in fsets.pop():
                by_factor.setdefault(f, []).append(n + 1)
        out["Supports"] = sup
        for f, lst in by_factor.items():
            lid = len(loads) + 1
            loads[str(lid)] = {"name": "SupportMotion", "attributes": {"list": lst}}
            lids.append(lid)
            facs.append(f)
    out["Loads"] = loads
    out["Combinations"] = {"1": {"name": "Combo", "attributes": {"load": lids, "factor": facs}}}
    out["Simulations"] = {"combo": 1, "attributes": {
        "solver": {"name": "EIGEN", "update": "OFF"},
        "algorithm": {"name": algorithm, "cnvgtol": cnvgtol, "nstep": nmax, "cnvgtest": cnvgtest},
        "integrator": {"name": integrator, "dt": dt, "mtol": 1e-12, "ktol": 1e-12, "ftol": 1e-12},
        "analysis": {"name": "DYNAMIC", "nt": nt}}}
    out["Recorders"] = {str(i + 1): {"name": "NODE", "file": f, "resp": r, "ndps": int(p), "nsamp": 1,
                                     "list": [int(n) for n in lst]}
                        for i, (f, r, lst, p) in enumerate(recorders)}
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    return np.arange(1, nn + 1)


def _fmt(v: float, ndps: int) -> str:
    return f"{v:.{ndps}e}"


def write_node_recorder(path: str, rec: NodeRecorder, node_tags, totaldof, ntotal: int, nt: int,
                        values: np.ndarray):
    """values[k-1, i, c]: response of component c of the recorder's node i after step k;
    node_tags / totaldof list every node the recorder may name."""
    totaldof = np.asarray(totaldof)
    ndof = totaldof.shape[1]
    tags = list(node_tags)
    rows = [tags.index(t) for t in rec.tags]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write(f"{len(rec.tags)} {len(rec.tags) * ndof} {ntotal} {nt}\n")
        for t, r in zip(rec.tags, rows):
            f.write(f"{t} {ndof}" + "".join(f" {int(d)}" for d in totaldof[r]) + "\n")
        for k in range(values.shape[0]):
            f.write("".join(_fmt(float(v), rec.ndps) + " " for v in values[k].reshape(-1)) + "\n")


def read_node_recorder(path: str):
    """Parses a NODE recorder file: (tags, totaldof [n, ndof], ntotal, nt, values [rows, n*ndof])."""
    with open(path) as f:
        lines = f.read().splitlines()
    nids, ndofs, ntotal, nt = (int(v) for v in lines[0].split())
    tags, tot = [], []
    for line in lines[1:1 + nids]:
        p = line.split()
        tags.append(int(p[0]))
        tot.append([int(v) for v in p[2:2 + int(p[1])]])
    vals = np.array([[float(v) for v in line.split()] for line in lines[1 + nids:] if line.strip()])
    return tags, np.asarray(tot), ntotal, nt, vals.reshape(-1, ndofs) if vals.size else np.zeros((0, ndofs))


def run(svl: SvlRun, out_dir: str | None = None, json_path: str | None = None, cg_rtol: float = 1e-12,
        learned: bool = False) -> dict:
    """Runs the simulation of a run file on the device (DynamicAnalysis k = 1..nt-1) and writes
    its NODE recorders into out_dir (default: SVL's <run dir>/../Solution/<combo>/). Returns
    {file: values}. VEL and ACCEL are formed from the recorded displacement history by the
    integrator's own update formulas (CentralDifference.cpp:141-148, ExtendedNewmarkBeta.cpp:
    141-150); REACTION from the device at the fixed mesh nodes (svlg_reactions) and as -eta V
    at dashpot partner nodes, 0 elsewhere; the NODE recorder is written for every step (SVL
    writes it on every call)."""
    from .integrators import CentralDifference, DeviceModel, ExtendedNewmarkBeta, LearnedPreconditioner, NewmarkBeta
    m = svl.model
    nd = m.ndim
    fd = m.free_dof()
    tags = list(svl.node_tags)
    mesh = set(tags)
    needed = {t for r in svl.recorders for t in r.tags if t in mesh}
    for r in svl.recorders:  # a partner's reaction is -eta V of its mesh node
        if r.response == "REACTION":
            for t in r.tags:
                for (n, _, _) in svl.partners.get(t, (None, []))[1]:
                    needed.add(tags[n])
    unknown = [t for r in svl.recorders for t in r.tags if t not in mesh and t not in svl.partners]
    if unknown:
        raise ValueError(f"recorder nodes {unknown[:5]} are not in the model")
    needed = sorted(needed)
    rec_dofs, where = [], {}
    moving = {(int(n), int(c)) for n, c in zip(m.sm_node, m.sm_dir)}
    for t in needed:
        for c in range(nd):
            f = int(fd[tags.index(t)][c])
            if f >= 0:
                where[(t, c)] = len(rec_dofs)
                rec_dofs.append(f)
            elif (tags.index(t), c) in moving:  # a restrained DOF with support motion: total DOF code
                where[(t, c)] = len(rec_dofs)
                rec_dofs.append(-1 - (tags.index(t) * nd + c))
    # Reactions are set only at fixed nodes (DynamicAnalysis.cpp:140-150); elsewhere 0.
    fixed_mesh = sorted({t for r in svl.recorders if r.response == "REACTION" for t in r.tags
                         if t in mesh and bool(m.restrained[tags.index(t)].any())})
    dm = DeviceModel(m)
    steps = svl.nt - 1
    if svl.integrator == "CENTRALDIFFERENCE":
        integ = CentralDifference(dm, svl.dt)
    elif svl.algorithm == "NEWTON":
        if svl.integrator != "NEWMARK":
            raise ValueError("the device Newton path is NewmarkBeta's (consistent residual)")
        integ = NewmarkBeta(dm, svl.dt, cnvgtol=svl.cnvgtol, nmax=svl.nmax, cnvgtest=svl.cnvgtest, rtol=cg_rtol)
    elif svl.integrator in ("NEWMARK", "EXTENDEDNEWMARK"):
        integ = ExtendedNewmarkBeta(dm, svl.dt, rtol=cg_rtol)
    else:
        dm.close()
        raise ValueError(f"integrator {svl.integrator} is not on the device path")
    integ.Initialize()
    if learned and not isinstance(integ, CentralDifference):
        LearnedPreconditioner(integ, channels=1, layers=8, mfma=-1).use(True)
    Rm = np.zeros((steps, len(fixed_mesh), nd))
    if fixed_mesh:
        # Step by step: the reactions of each step's state.
        qn = [tags.index(t) for t in fixed_mesh]
        rows = []
        for k in range(1, steps + 1):
            h = integ.advance(k, 1, rec_dofs) if rec_dofs else None
            rows.append(np.asarray(h).reshape(-1) if h is not None else np.zeros(0))
            Rm[k - 1] = dm.reactions(qn, k)
        U = np.asarray(rows).reshape(steps, len(rec_dofs))
    else:
        U = integ.advance(1, steps, rec_dofs) if rec_dofs else (integ.advance(1, steps) or np.zeros((steps, 0)))
    dm.close()
    U = np.asarray(U).reshape(steps, len(rec_dofs))
    Up = np.vstack([np.zeros((1, U.shape[1])), U[:-1]])
    dt = svl.dt
    if svl.integrator == "CENTRALDIFFERENCE":
        Upp = np.vstack([np.zeros((2, U.shape[1])), U[:-2]])[:steps]
        dU = U - Up
        V = 1.0 / 2.0 / dt * ((Up + dU) - Upp)
        A = 1.0 / dt / dt * ((dU - Up) + Upp)
    else:
        V, A = np.zeros_like(U), np.zeros_like(U)
        v, a = np.zeros(U.shape[1]), np.zeros(U.shape[1])
        # V and A follow the free increment alone (ExtendedNewmarkBeta.cpp:134-150): 0 at the
        # support-motion DOFs.
        free_col = np.asarray([r >= 0 for r in rec_dofs], dtype=bool)
        for k in range(steps):
            du = np.where(free_col, U[k] - Up[k], 0.0)
            a = 4.0 / dt / dt * du - 4.0 / dt * v - a
            v = 2.0 / dt * du - v
            V[k], A[k] = v, a
    resp = {"DISP": U, "VEL": V, "ACCEL": A}
    base = out_dir
    if base is None:
        root = os.path.dirname(os.path.abspath(json_path)) if json_path else "."
        base = os.path.join(root, "..", "Solution", svl.combo_name)
    all_tags = tags + list(svl.partners)
    all_tot = np.vstack([svl.totaldof] + [np.asarray([v[0]]) for v in svl.partners.values()])
    out = {}
    for r in svl.recorders:
        vals = np.zeros((steps, len(r.tags), nd))
        for i, t in enumerate(r.tags):
            if r.response == "REACTION":
                if t in fixed_mesh:
                    vals[:, i, :] = Rm[:, fixed_mesh.index(t), :]
                elif t in svl.partners:
                    for (n, c, eta) in svl.partners[t][1]:
                        j = where.get((tags[n], c))
                        if j is not None:
                            vals[:, i, c] += -eta * V[:, j]
                continue
            if t not in mesh:
                continue  # restrained partner node: its motion is zero
            for c in range(nd):
                j = where.get((t, c))
                if j is not None:
                    vals[:, i, c] = resp[r.response][:, j]
        p = os.path.join(base, r.file)
        write_node_recorder(p, r, all_tags, all_tot, svl.ntotal, svl.nt, vals)
        out[p] = vals
    return out


def main(argv=None):
    """Command line: ``python -m svl_amd.svlfile RUN.json [--out DIR] [--learned]`` runs an
    SVL run file on the GPU and writes its NODE recorders (default: SVL's
    <run dir>/../Solution/<combo>/)."""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("file")
    ap.add_argument("--out", default=None)
    ap.add_argument("--cg-rtol", type=float, default=1e-12)
    ap.add_argument("--learned", action="store_true", help="learned preconditioner for the implicit solves")
    a = ap.parse_args(argv)
    from .pmlfile import is_pml_run, read_pml_run, run_pml
    if is_pml_run(a.file):
        r = read_pml_run(a.file)
        files = run_pml(r, out_dir=a.out, json_path=a.file)
    else:
        r = read_run(a.file)
        files = run(r, out_dir=a.out, json_path=a.file, cg_rtol=a.cg_rtol, learned=a.learned)
    for p in files:
        print(p)
    for s in r.skipped_recorders:
        print(f"skipped recorder {s.get('name')} ({s.get('file')}): not written by the device path")


if __name__ == "__main__":
    main()
