"""Perfectly matched layers: mixed models of PML2DQuad4 / lin2DQuad4 (2D) or PML3DHexa8 /
lin3DHexa8 (3D) elements tied by EQUAL constraints, stepped on the device by
ExtendedNewmarkBeta + Linear (libsvlg's svlg_pml_* entry points).

A PML node carries five unknowns in 2D (ux, uy and the stress-like xx, yy, xy;
04-Elements/06-Quadrilateral/PML2DQuad4.cpp:11-12) and nine in 3D (ux, uy, uz, xx, yy, zz,
xy, yz, xz; 04-Elements/10-Hexahedron/PML3DHexa8.cpp), a regular node ndim. The free numbering
follows SVL's Mesh (06-Mesh/Mesh.cpp:328-383): nodes in ascending order, restrained unknowns
and EQUAL slaves skipped, a slave taking its master's free number times the constraint factor.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi

LIN2DQUAD4, PML2DQUAD4, LIN3DHEXA8, PML3DHEXA8 = 0, 1, 2, 3
NEN = {LIN2DQUAD4: 4, PML2DQUAD4: 4, LIN3DHEXA8: 8, PML3DHEXA8: 8}


class MixedDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("nnodes", C.c_int32),
        ("coords", capi.f64p),
        ("dof_ptr", capi.i32p),
        ("dof_free", capi.i32p),
        ("dof_factor", capi.f64p),
        ("nelem", C.c_int32),
        ("elem_class", capi.i32p),
        ("conn", capi.i32p),
        ("E", capi.f64p),
        ("nu", capi.f64p),
        ("rho", capi.f64p),
        ("th", capi.f64p),
        ("pml", capi.f64p),
        ("lumped", C.c_int32),
        ("nload", C.c_int32),
        ("load_nt", C.c_int32),
        ("load_dof", capi.i32p),
        ("load_factor", capi.f64p),
        ("load_series", capi.f64p),
    ]


@dataclass
class PmlMesh:
    """Nodes with 2 or 5 (2D) or 3 or 9 (3D) unknowns; elements of class LIN2DQUAD4 /
    PML2DQUAD4 / LIN3DHEXA8 / PML3DHEXA8 with E, nu, rho, thickness (2D) and, for PML
    elements, (m, L, R, x0, y0, z0, nx, ny, nz) in the order of the PML constructors (z0, nz
    unused in 2D); EQUAL constraints (slave node, slave unknown, master node, master unknown,
    factor); point loads (node, unknown, factor, series)."""
    coords: np.ndarray                       # [nnodes, ndim]
    ndof: np.ndarray                         # [nnodes]
    fixed: list                              # per node: booleans, one per unknown
    elem_class: np.ndarray                   # [nelem]
    conn: np.ndarray                         # [nelem, 4 or 8]
    E: np.ndarray
    nu: np.ndarray
    rho: np.ndarray
    th: np.ndarray
    pml: np.ndarray                          # [nelem, 9]
    equal: list = field(default_factory=list)
    loads: list = field(default_factory=list)
    lumped: bool = False

    def __post_init__(self):
        self.coords = np.ascontiguousarray(self.coords, np.float64)
        self.ndof = np.ascontiguousarray(self.ndof, np.int32)
        self.elem_class = np.ascontiguousarray(self.elem_class, np.int32)
        self.conn = np.ascontiguousarray(self.conn, np.int32)
        ne = len(self.conn)
        for name in ("E", "nu", "rho", "th"):
            setattr(self, name, np.ascontiguousarray(np.broadcast_to(getattr(self, name), (ne,)), np.float64))
        self.pml = np.ascontiguousarray(np.broadcast_to(self.pml, (ne, 9)), np.float64)

    @property
    def ndim(self) -> int:
        return int(self.coords.shape[1])

    def numbering(self):
        """(dof_ptr [nnodes+1], dof_free [ntotal], dof_factor [ntotal], nfree)."""
        ptr = np.concatenate([[0], np.cumsum(self.ndof)]).astype(np.int32)
        slaves = {(s, a): (m, b, f) for s, a, m, b, f in self.equal}
        free = np.full(int(ptr[-1]), -1, np.int32)
        fac = np.ones(int(ptr[-1]))
        nfree = 0
        for n in range(len(self.ndof)):
            for a in range(int(self.ndof[n])):
                if (n, a) in slaves or self.fixed[n][a]:
                    continue
                free[ptr[n] + a] = nfree
                nfree += 1
        for (s, a), (m, b, f) in slaves.items():
            free[ptr[s] + a] = free[ptr[m] + b]
            fac[ptr[s] + a] = f
        return ptr, free, fac, nfree

    def total_dof(self, node: int, unknown: int) -> int:
        return int(np.concatenate([[0], np.cumsum(self.ndof)])[node] + unknown)


class PmlNewmark:
    """ExtendedNewmarkBeta + Linear on a PmlMesh (10-Integrators/ExtendedNewmarkBeta.cpp;
    the 2D element has no history matrix, so G and Ubar vanish). The effective matrix is
    assembled once on the device and solved by MINRES to `rtol` every step."""

    def __init__(self, mesh: PmlMesh, dt: float, rtol: float = 1e-14, maxit: int = 4000):
        self.mesh = mesh
        ptr, free, fac, nfree = mesh.numbering()
        self.dof_ptr, self.dof_free, self.dof_factor, self.nfree = ptr, free, fac, nfree
        nt = max((len(s) for _, _, _, s in mesh.loads), default=0)
        series = np.zeros((len(mesh.loads), max(nt, 1)))
        for i, (_, _, _, s) in enumerate(mesh.loads):
            series[i, :len(s)] = s
        ldof = np.array([ptr[n] + a for n, a, _, _ in mesh.loads], np.int32)
        lfac = np.array([f for _, _, f, _ in mesh.loads], np.float64)
        keep = [ptr, free, fac, series, ldof, lfac]
        d = MixedDesc()
        d.ndim = mesh.ndim
        d.nnodes = len(mesh.ndof)
        d.coords = mesh.coords.ctypes.data_as(capi.f64p)
        d.dof_ptr = capi.iptr(ptr)
        d.dof_free = capi.iptr(free)
        d.dof_factor = capi.fptr(fac)
        d.nelem = len(mesh.conn)
        d.elem_class = capi.iptr(mesh.elem_class)
        d.conn = capi.iptr(mesh.conn)
        d.E, d.nu, d.rho, d.th = (capi.fptr(a) for a in (mesh.E, mesh.nu, mesh.rho, mesh.th))
        d.pml = capi.fptr(mesh.pml)
        d.lumped = int(bool(mesh.lumped))
        d.nload = len(mesh.loads)
        d.load_nt = int(series.shape[1]) if mesh.loads else 0
        d.load_dof = capi.iptr(ldof)
        d.load_factor = capi.fptr(lfac)
        d.load_series = capi.fptr(series)
        h = C.c_void_p()
        capi.check(capi.lib().svlg_pml_create(C.byref(d), float(dt), float(rtol), int(maxit), C.byref(h)),
                   "svlg_pml_create")
        del keep
        self.h = h
        self.dt = dt

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            capi.lib().svlg_pml_destroy(h)
            self.h = None

    def info(self) -> dict:
        v = np.zeros(4, np.int64)
        capi.check(capi.lib().svlg_pml_info(self.h, v.ctypes.data_as(C.POINTER(C.c_int64))), "svlg_pml_info")
        return dict(nfree=int(v[0]), ntotal=int(v[1]), nnz=int(v[2]), contributions=int(v[3]))

    def advance(self, k0: int, nsteps: int, rec=()):
        """Steps k0 .. k0+nsteps-1. `rec` lists (node, unknown) pairs; returns (U, V, A) histories
        [nsteps, len(rec)] of those total DOFs and the MINRES iterations per step."""
        rd = np.array([self.dof_ptr[n] + a for n, a in rec], np.int32)
        hu, hv, ha = (np.zeros((nsteps, rd.size)) for _ in range(3))
        its = np.zeros(max(nsteps, 1), np.int32)
        capi.check(capi.lib().svlg_pml_steps(self.h, int(k0), int(nsteps), int(rd.size), capi.iptr(rd),
                                             capi.fptr(hu), capi.fptr(hv), capi.fptr(ha), capi.iptr(its)),
                   "svlg_pml_steps")
        return hu, hv, ha, its[:nsteps]

    def state(self):
        U, V, A = (np.zeros(self.nfree) for _ in range(3))
        capi.check(capi.lib().svlg_pml_get_state(self.h, capi.fptr(U), capi.fptr(V), capi.fptr(A)),
                   "svlg_pml_get_state")
        return U, V, A

    def GetDisplacements(self):
        return self.state()[0]

    def set_graph(self, on: bool):
        """Replay the MINRES iterations from a captured hipGraph (default) or launch them eagerly."""
        capi.check(capi.lib().svlg_pml_set_graph(self.h, int(bool(on))), "svlg_pml_set_graph")

    def GetPMLHistoryVector(self):
        """Ubar over the free DOFs (ExtendedNewmarkBeta's integrated displacement; 0 in 2D)."""
        out = np.zeros(self.nfree)
        capi.check(capi.lib().svlg_pml_get_history(self.h, capi.fptr(out)), "svlg_pml_get_history")
        return out

    def GetVelocities(self):
        return self.state()[1]

    def GetAccelerations(self):
        return self.state()[2]


def soil_column(n_reg: int, n_pml: int, load_series, E=5e7, nu=0.25, rho=2000.0, m=2.0, L=None, R=1e-5,
                lumped=False) -> PmlMesh:
    """A 1 x (n_pml + n_reg) column of unit squares: n_pml PML2DQuad4 elements at the bottom
    (normal (0, -1), the layer's top at y = n_pml) under n_reg lin2DQuad4 elements, ux
    restrained everywhere and uy at the base, the coincident interface nodes tied by EQUAL
    constraints on uy (PML node slave, regular node master), a vertical load of 1e5 x series
    at both top nodes. With n_pml = 10, n_reg = 90 this is SVL's validation case F11."""
    L = float(n_pml if L is None else L)
    coords, ndof, fixed, cls, conn = [], [], [], [], []
    for j in range(n_pml + 1):
        for i in range(2):
            coords.append([float(i), float(j)])
            ndof.append(5)
            fixed.append([True, j == 0, False, False, False])
    for j in range(n_pml):
        cls.append(PML2DQUAD4)
        conn.append([2 * j, 2 * j + 1, 2 * j + 3, 2 * j + 2])
    r0 = len(coords)
    for j in range(n_reg + 1):
        for i in range(2):
            coords.append([float(i), float(n_pml + j)])
            ndof.append(2)
            fixed.append([True, False])
    for j in range(n_reg):
        b = r0 + 2 * j
        cls.append(LIN2DQUAD4)
        conn.append([b, b + 1, b + 3, b + 2])
    equal = [(2 * n_pml, 1, r0, 1, 1.0), (2 * n_pml + 1, 1, r0 + 1, 1, 1.0)]
    top = len(coords) - 2
    s = 1e5 * np.asarray(load_series, np.float64)
    pml = np.array([m, L, R, 0.5, float(n_pml), 0.0, 0.0, -1.0, 0.0])
    return PmlMesh(np.array(coords), np.array(ndof), fixed, np.array(cls), np.array(conn), E, nu, rho, 1.0, pml,
                   equal=equal, loads=[(top, 1, 1.0, s), (top + 1, 1, 1.0, s)], lumped=lumped)


def halfspace_box(nx: int, ny: int, npml: int, h: float = 1.0, E=5e7, nu=0.25, rho=2000.0, m=2.0, R=1e-5,
                  lumped=False, load_series=None) -> PmlMesh:
    """A 2D half-space: an nx x ny block of lin2DQuad4 elements (size h) wrapped on its left,
    right and bottom sides by PML2DQuad4 layers npml elements thick (corner elements stretched
    in both directions). PML nodes that coincide with regular ones are separate nodes tied by
    EQUAL constraints on ux and uy; the outer PML boundary is fixed in ux and uy. A vertical
    point load of 1e5 x series acts at the top centre."""
    L = npml * h
    X0, Y0 = npml * h, npml * h             # lower-left corner of the regular block
    NX, NY = nx + 2 * npml, ny + npml       # element grid of the whole domain
    coords, ndof, fixed = [], [], []
    reg_id, pml_id = {}, {}

    def in_reg_elem(i, j):
        return npml <= i < npml + nx and j >= npml

    is_pml_node, is_reg_node = set(), set()
    for j in range(NY):
        for i in range(NX):
            corners = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
            (is_reg_node if in_reg_elem(i, j) else is_pml_node).update(corners)
    for j in range(NY + 1):
        for i in range(NX + 1):
            x, y = i * h, j * h
            if (i, j) in is_pml_node:
                pml_id[(i, j)] = len(coords)
                coords.append([x, y])
                ndof.append(5)
                outer = i == 0 or i == NX or j == 0
                fixed.append([outer, outer, False, False, False])
            if (i, j) in is_reg_node:
                reg_id[(i, j)] = len(coords)
                coords.append([x, y])
                ndof.append(2)
                fixed.append([False, False])
    cls, conn, pml = [], [], []
    for j in range(NY):
        for i in range(NX):
            corners = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
            if in_reg_elem(i, j):
                cls.append(LIN2DQUAD4)
                conn.append([reg_id[c] for c in corners])
                pml.append([m, L, R, 0, 0, 0, 0, 0, 0])
            else:
                cls.append(PML2DQUAD4)
                conn.append([pml_id[c] for c in corners])
                xc, yc = (i + 0.5) * h, (j + 0.5) * h
                nxv = -1.0 if xc < X0 else (1.0 if xc > X0 + nx * h else 0.0)
                nyv = -1.0 if yc < Y0 else 0.0
                x0 = X0 if nxv < 0 else X0 + nx * h
                pml.append([m, L, R, x0, Y0, 0.0, nxv, nyv, 0.0])
    equal = []
    for key, r in reg_id.items():
        if key in pml_id:
            equal += [(pml_id[key], 0, r, 0, 1.0), (pml_id[key], 1, r, 1, 1.0)]
    loads = []
    if load_series is not None:
        top = reg_id[(npml + nx // 2, NY)]
        loads = [(top, 1, 1.0, 1e5 * np.asarray(load_series, np.float64))]
    return PmlMesh(np.array(coords), np.array(ndof), fixed, np.array(cls), np.array(conn), E, nu, rho, 1.0,
                   np.array(pml), equal=equal, loads=loads, lumped=lumped)


def soil_rod3d(n_reg: int, n_pml: int, load_series, E=5e7, nu=0.25, rho=2000.0, m=2.0, L=None, R=1e-5,
               lumped=False) -> PmlMesh:
    """A 1 x 1 x (n_pml + n_reg) rod of unit cubes: n_pml PML3DHexa8 elements at the bottom
    (normal (0, 0, -1), the layer's top at z = n_pml) under n_reg lin3DHexa8 elements, ux and
    uy restrained everywhere and uz at the base, the coincident interface nodes tied by EQUAL
    constraints on uz, a vertical load of 0.5e5 x series at the four top nodes. With
    n_pml = 10, n_reg = 90 this is SVL's validation case J12."""
    L = float(n_pml if L is None else L)
    coords, ndof, fixed, cls, conn = [], [], [], [], []

    def level(z, nun, fix):
        for y in (0.0, 1.0):
            for x in (0.0, 1.0):
                coords.append([x, y, z])
                ndof.append(nun)
                fixed.append(fix)

    def hexa(b):
        return [b, b + 1, b + 3, b + 2, b + 4, b + 5, b + 7, b + 6]

    for k in range(n_pml + 1):
        level(float(k), 9, [True, True, k == 0] + [False] * 6)
    for k in range(n_pml):
        cls.append(PML3DHEXA8)
        conn.append(hexa(4 * k))
    r0 = len(coords)
    for k in range(n_reg + 1):
        level(float(n_pml + k), 3, [True, True, False])
    for k in range(n_reg):
        cls.append(LIN3DHEXA8)
        conn.append(hexa(r0 + 4 * k))
    equal = [(4 * n_pml + i, 2, r0 + i, 2, 1.0) for i in range(4)]
    top = len(coords) - 4
    s = 0.5e5 * np.asarray(load_series, np.float64)
    pml = np.array([m, L, R, 0.5, 0.5, float(n_pml), 0.0, 0.0, -1.0])
    return PmlMesh(np.array(coords), np.array(ndof), fixed, np.array(cls), np.array(conn), E, nu, rho, 1.0, pml,
                   equal=equal, loads=[(top + i, 2, 1.0, s) for i in range(4)], lumped=lumped)


def halfspace_box3d(nx: int, ny: int, nz: int, npml: int, h: float = 1.0, E=5e7, nu=0.25, rho=2000.0, m=2.0,
                    R=1e-5, lumped=False, load_series=None) -> PmlMesh:
    """A 3D half-space: an nx x ny x nz block of lin3DHexa8 elements (size h) wrapped on its
    four sides and its bottom by PML3DHexa8 layers npml elements thick (edge and corner
    elements stretched in two or three directions, which gives them a history matrix). PML
    nodes coinciding with regular ones are separate nodes tied by EQUAL constraints on ux, uy,
    uz; the outer PML boundary is fixed. A vertical point load of 1e5 x series acts at the
    top node nearest the centre."""
    L = npml * h
    X0 = Y0 = Z0 = npml * h
    NX, NY, NZ = nx + 2 * npml, ny + 2 * npml, nz + npml

    def is_reg(i, j, k):
        return npml <= i < npml + nx and npml <= j < npml + ny and k >= npml

    def corners(i, j, k):
        return [(i, j, k), (i + 1, j, k), (i + 1, j + 1, k), (i, j + 1, k),
                (i, j, k + 1), (i + 1, j, k + 1), (i + 1, j + 1, k + 1), (i, j + 1, k + 1)]

    pml_pts, reg_pts = set(), set()
    for k in range(NZ):
        for j in range(NY):
            for i in range(NX):
                (reg_pts if is_reg(i, j, k) else pml_pts).update(corners(i, j, k))
    coords, ndof, fixed, reg_id, pml_id = [], [], [], {}, {}
    for k in range(NZ + 1):
        for j in range(NY + 1):
            for i in range(NX + 1):
                pt = (i, j, k)
                if pt in pml_pts:
                    pml_id[pt] = len(coords)
                    coords.append([i * h, j * h, k * h])
                    ndof.append(9)
                    outer = i in (0, NX) or j in (0, NY) or k == 0
                    fixed.append([outer] * 3 + [False] * 6)
                if pt in reg_pts:
                    reg_id[pt] = len(coords)
                    coords.append([i * h, j * h, k * h])
                    ndof.append(3)
                    fixed.append([False] * 3)
    cls, conn, pml = [], [], []
    for k in range(NZ):
        for j in range(NY):
            for i in range(NX):
                if is_reg(i, j, k):
                    cls.append(LIN3DHEXA8)
                    conn.append([reg_id[c] for c in corners(i, j, k)])
                    pml.append([m, L, R] + [0.0] * 6)
                    continue
                cls.append(PML3DHEXA8)
                conn.append([pml_id[c] for c in corners(i, j, k)])
                xc, yc, zc = (i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h
                nxv = -1.0 if xc < X0 else (1.0 if xc > X0 + nx * h else 0.0)
                nyv = -1.0 if yc < Y0 else (1.0 if yc > Y0 + ny * h else 0.0)
                nzv = -1.0 if zc < Z0 else 0.0
                pml.append([m, L, R, X0 if nxv <= 0 else X0 + nx * h, Y0 if nyv <= 0 else Y0 + ny * h, Z0,
                            nxv, nyv, nzv])
    equal = []
    for pt, r in reg_id.items():
        if pt in pml_id:
            equal += [(pml_id[pt], a, r, a, 1.0) for a in range(3)]
    loads = []
    if load_series is not None:
        top = reg_id[(npml + nx // 2, npml + ny // 2, NZ)]
        loads = [(top, 2, 1.0, 1e5 * np.asarray(load_series, np.float64))]
    return PmlMesh(np.array(coords), np.array(ndof), fixed, np.array(cls), np.array(conn), E, nu, rho, 1.0,
                   np.array(pml), equal=equal, loads=loads, lumped=lumped)
