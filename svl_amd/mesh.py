"""Finite-element model description in SeismoVLAB's pre-processor conventions.

This module is the host-side mirror of the parts of SVL's Python pre-processor that
decide *what* the dynamic hot path computes on:

* structured mesh generation with SVL's node and element tag order
  (01-Pre_Process/Method/Builder.py:57-242 ``makeDomainVolume``,
  Builder.py:244-417 ``makeDomainArea``);
* boundary node sets (Builder.py:669-773 ``Find2DBoundaries``/``Find3DBoundaries``);
* Lysmer dashpots as ZeroLength1D + Viscous1DLinear elements tied to new, fully
  restrained copies of the boundary nodes (Builder.py:1086-1131 ``setLysmerDomain``);
* the Plain degree-of-freedom numbering (01-Pre_Process/Core/Numberer.py:185-233);
* the Ricker velocity pulse (01-Pre_Process/Core/PlaneWave.py:222-223).

A :class:`Model` holds plain numpy arrays in ascending-tag order; the device library
(``libsvlg``) and the CPU oracle both consume them through ``svlg_mesh_desc``.

Deviation kept on purpose: SVL's ``setLysmerDomain`` iterates CPython ``set`` objects,
so the tags it gives to the new Lysmer nodes and dashpot elements depend on hash order.
Here boundary sets are visited in the order the caller names them and, inside each set,
in ascending node tag. Free-DOF numbering is unaffected (Lysmer nodes are restrained and
tagged after every mesh node).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

ELEM_QUAD4 = 0
ELEM_HEXA8 = 1

MAT_ELASTIC = 0
MAT_J2 = 1


@dataclass
class Model:
    """A continuum mesh plus SVL-style dashpots, loads and damping, tag-ordered."""

    ndim: int
    coords: np.ndarray                  # [nnodes, ndim] float64, mesh nodes in ascending tag
    node_tags: np.ndarray               # [nnodes] int64 SVL node tags (ascending)
    elem_type: int                      # ELEM_QUAD4 / ELEM_HEXA8
    conn: np.ndarray                    # [nelem, nen] int32, 0-based mesh-node indices
    E: np.ndarray                       # [nelem]
    nu: np.ndarray                      # [nelem]
    rho: np.ndarray                     # [nelem]
    thickness: float = 1.0
    restrained: np.ndarray | None = None    # [nnodes, ndim] bool
    mat_type: np.ndarray | None = None      # [nelem] int32 (MAT_ELASTIC / MAT_J2)
    # J2 parameters per element (Plastic3DJ2 / PlasticPlaneStrainJ2): K, G, sigma_y, H, beta
    j2: np.ndarray | None = None            # [nelem, 5]
    ray_am: np.ndarray | None = None        # [nelem] Rayleigh mass coefficient
    ray_ak: np.ndarray | None = None        # [nelem] Rayleigh stiffness coefficient
    # Dashpots: ZeroLength1D(Viscous1DLinear) from mesh node to a restrained copy.
    dash_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    dash_dir: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    dash_eta: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    # Point loads (POINTLOAD_CONCENTRATED_DYNAMIC): node, direction vector, factor, series.
    load_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    load_dir: np.ndarray = field(default_factory=lambda: np.zeros((0, 1), np.float64))
    load_factor: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    load_series: np.ndarray = field(default_factory=lambda: np.zeros((0, 1), np.float64))
    massform_lumped: bool = True
    name: str = "model"
    # Free-DOF numbers of a run file ([nnodes, ndim], -1 restrained); None: the Plain scheme.
    free_dof_override: np.ndarray | None = None
    # Domain reduction (ElementLoad GENERALWAVE): DRM elements, their nodes with the exterior
    # flag, and per node the signal rows [nt, 3*ndim] (U, V, A components, as in SVL's files).
    drm_elem: np.ndarray | None = None
    drm_node: np.ndarray | None = None
    drm_cond: np.ndarray | None = None
    drm_signal: np.ndarray | None = None     # [n_drm_node, nt, 3*ndim]
    drm_factor: float = 1.0
    # Support motion (Supports + POINTLOAD_SUPPORT_MOTION of the combination): per entry a
    # restrained node, its direction, the load's combination factor and the prescribed
    # displacement series g (read at step k as g[k], or g[0] past its end, as Node::
    # GetSupportMotion does).
    sm_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    sm_dir: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    sm_factor: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    sm_series: list = field(default_factory=list)
    sm_offset: int = 0      # 1: step k applies the increment of sample k-1 (NewmarkBeta)

    def __post_init__(self):
        self.coords = np.ascontiguousarray(self.coords, dtype=np.float64)
        self.conn = np.ascontiguousarray(self.conn, dtype=np.int32)
        n = self.coords.shape[0]
        ne = self.conn.shape[0]
        self.E = np.array(np.broadcast_to(self.E, (ne,)), dtype=np.float64)
        self.nu = np.array(np.broadcast_to(self.nu, (ne,)), dtype=np.float64)
        self.rho = np.array(np.broadcast_to(self.rho, (ne,)), dtype=np.float64)
        if self.restrained is None:
            self.restrained = np.zeros((n, self.ndim), dtype=bool)
        self.restrained = np.ascontiguousarray(self.restrained, dtype=bool)
        if self.mat_type is None:
            self.mat_type = np.zeros(ne, dtype=np.int32)
        self.mat_type = np.ascontiguousarray(self.mat_type, dtype=np.int32)
        if self.j2 is None:
            self.j2 = np.zeros((ne, 5), dtype=np.float64)
        self.j2 = np.ascontiguousarray(self.j2, dtype=np.float64)
        if self.ray_am is None:
            self.ray_am = np.zeros(ne)
        if self.ray_ak is None:
            self.ray_ak = np.zeros(ne)
        self.ray_am = np.array(np.broadcast_to(self.ray_am, (ne,)), dtype=np.float64)
        self.ray_ak = np.array(np.broadcast_to(self.ray_ak, (ne,)), dtype=np.float64)
        self.dash_node = np.ascontiguousarray(self.dash_node, dtype=np.int32)
        self.dash_dir = np.ascontiguousarray(self.dash_dir, dtype=np.int32)
        self.dash_eta = np.ascontiguousarray(self.dash_eta, dtype=np.float64)
        self.load_node = np.ascontiguousarray(self.load_node, dtype=np.int32)
        self.load_dir = np.ascontiguousarray(self.load_dir, dtype=np.float64).reshape(-1, self.ndim) \
            if self.load_node.size else np.zeros((0, self.ndim))
        self.load_factor = np.ascontiguousarray(self.load_factor, dtype=np.float64)
        if self.load_node.size:
            self.load_series = np.ascontiguousarray(self.load_series, dtype=np.float64).reshape(self.load_node.size, -1)
        else:
            self.load_series = np.zeros((0, 1))
        self.load_dir = np.ascontiguousarray(self.load_dir)
        self.load_series = np.ascontiguousarray(self.load_series)

    def add_support_motion(self, node: int, direction: int, series, factor: float = 1.0):
        """Prescribed displacement history of a restrained DOF (Supports TIMESERIES / CONSTANT
        activated by a SupportMotion load with the given combination factor)."""
        if not self.restrained[node, direction]:
            raise ValueError("support motion needs a restrained DOF")
        self.sm_node = np.append(self.sm_node, np.int32(node)).astype(np.int32)
        self.sm_dir = np.append(self.sm_dir, np.int32(direction)).astype(np.int32)
        self.sm_factor = np.append(self.sm_factor, float(factor)).astype(np.float64)
        self.sm_series = list(self.sm_series) + [np.ascontiguousarray(series, dtype=np.float64).reshape(-1)]

    def set_j2(self, K, G, sigma_y, H, beta, elems=None):
        """Makes elements J2 (Plastic3DJ2 / PlasticPlaneStrainJ2) with bulk modulus K, shear
        modulus G, yield stress sigma_y, hardening modulus H and mixing beta (1 isotropic,
        0 kinematic); E and nu are set to the equivalent elastic constants."""
        idx = np.arange(self.nelem) if elems is None else np.asarray(elems)
        self.mat_type[idx] = MAT_J2
        self.j2[idx] = np.stack(np.broadcast_arrays(K, G, sigma_y, H, beta), axis=-1)
        K = np.broadcast_to(K, idx.shape)
        G = np.broadcast_to(G, idx.shape)
        self.E[idx] = 9.0 * K * G / (3.0 * K + G)
        self.nu[idx] = (3.0 * K - 2.0 * G) / (2.0 * (3.0 * K + G))

    # ------------------------------------------------------------------ sizes
    @property
    def nnodes(self) -> int:
        return self.coords.shape[0]

    @property
    def nelem(self) -> int:
        return self.conn.shape[0]

    @property
    def nen(self) -> int:
        return self.conn.shape[1]

    @property
    def nfree(self) -> int:
        return int((~self.restrained).sum())

    @property
    def n_lysmer_nodes(self) -> int:
        return int(np.unique(self.dash_node).size) if self.dash_node.size else 0

    @property
    def ntotal(self) -> int:
        """SVL's numberOfTotalDofs: mesh nodes plus the restrained Lysmer copies."""
        return (self.nnodes + self.n_lysmer_nodes) * self.ndim

    def free_dof(self) -> np.ndarray:
        """Plain-scheme free DOF index per (node, component); -1 where restrained.

        Follows 01-Pre_Process/Core/Numberer.py:207-228: nodes in ascending tag,
        components in order, restrained components skipped. Lysmer copies are tagged
        after every mesh node and are fully restrained, so they take no free index.
        """
        if self.free_dof_override is not None:
            return np.ascontiguousarray(self.free_dof_override, dtype=np.int32)
        free = np.full((self.nnodes, self.ndim), -1, dtype=np.int32)
        mask = ~self.restrained
        free[mask] = np.arange(int(mask.sum()), dtype=np.int32)
        return free

    def total_dof(self) -> np.ndarray:
        """Plain-scheme total DOF index per (node, component) for mesh nodes."""
        return np.arange(self.nnodes * self.ndim, dtype=np.int32).reshape(self.nnodes, self.ndim)

    def node_index(self, tag: int) -> int:
        idx = np.searchsorted(self.node_tags, tag)
        if idx >= self.node_tags.size or self.node_tags[idx] != tag:
            raise KeyError(f"node tag {tag} not in model")
        return int(idx)

    def nearest_node(self, xyz) -> int:
        d = np.linalg.norm(self.coords - np.asarray(xyz, dtype=np.float64)[None, :], axis=1)
        return int(np.argmin(d))

    def add_point_load(self, node: int, direction, series, factor: float = 1.0):
        """Adds a TIMESERIES PointLoad on one node (Driver.hpp:1512-1535, Load.cpp:69-71)."""
        d = np.asarray(direction, dtype=np.float64).reshape(1, self.ndim)
        s = np.asarray(series, dtype=np.float64).reshape(1, -1)
        if self.load_node.size:
            if s.shape[1] != self.load_series.shape[1]:
                raise ValueError("all load series of a model must have the same length")
            self.load_series = np.ascontiguousarray(np.vstack([self.load_series, s]))
            self.load_dir = np.ascontiguousarray(np.vstack([self.load_dir, d]))
        else:
            self.load_series = np.ascontiguousarray(s)
            self.load_dir = np.ascontiguousarray(d)
        self.load_node = np.append(self.load_node, np.int32(node)).astype(np.int32)
        self.load_factor = np.append(self.load_factor, float(factor)).astype(np.float64)

    def lame(self):
        """Per-element (lambda, mu) of the isotropic D matrix (Elastic3DLinear.cpp:19-28)."""
        E, nu = self.E, self.nu
        lam = E * nu / (1.0 - 2.0 * nu) / (1.0 + nu)
        mu = E / (2.0 * (1.0 + nu))
        return lam, mu

    def min_cfl_dt(self) -> float:
        """h_min / Vp_max over elements, h = shortest element edge along the grid."""
        lam, mu = self.lame()
        vp = np.sqrt((lam + 2.0 * mu) / self.rho)
        x = self.coords[self.conn]
        h = np.linalg.norm(x[:, 1] - x[:, 0], axis=1)
        if self.elem_type == ELEM_HEXA8:
            h = np.minimum(h, np.linalg.norm(x[:, 3] - x[:, 0], axis=1))
            h = np.minimum(h, np.linalg.norm(x[:, 4] - x[:, 0], axis=1))
        else:
            h = np.minimum(h, np.linalg.norm(x[:, 3] - x[:, 0], axis=1))
        return float(np.min(h / vp))


# ---------------------------------------------------------------------- builders
def make_domain_area(nx: int, ny: int, P0, P1, P2):
    """Quad4 grid with SVL's tag order (Builder.py:317-323 nodes, :372-379 elements).

    Returns (coords[nnodes,2], conn[nelem,4] 0-based, boundary dict of node-index arrays).
    Node tag t (1-based) = (nx+1)*j + i + 1 for grid point (i, j).
    """
    P0, P1, P2 = (np.asarray(p, dtype=np.float64) for p in (P0, P1, P2))
    DX = (P1 - P0) / nx
    DY = (P2 - P0) / ny
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = P0[None, :] + DX[None, :] * i.reshape(-1, 1) + DY[None, :] * j.reshape(-1, 1)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ii = ii.reshape(-1)
    jj = jj.reshape(-1)
    n1 = (nx + 1) * jj + ii
    conn = np.stack([n1, n1 + 1, n1 + (nx + 1) + 1, n1 + (nx + 1)], axis=1).astype(np.int32)
    # Find2DBoundaries (Builder.py:669-714) with TOL = DX/1000.
    tolx, toly = DX[0] / 1000.0, DY[1] / 1000.0
    top, bottom = P2[1] - toly, P0[1] + toly
    left, right = P0[0] + tolx, P1[0] - tolx
    x, y = coords[:, 0], coords[:, 1]
    bnd = {
        "right": np.nonzero(x - right > 0.0)[0],
        "left": np.nonzero((x - right <= 0.0) & (left - x > 0.0))[0],
        "top": np.nonzero(y - top > 0.0)[0],
        "bottom": np.nonzero((y - top <= 0.0) & (bottom - y > 0.0))[0],
    }
    return coords, conn, bnd


def make_domain_volume(nx: int, ny: int, nz: int, P0, P1, P2, P3):
    """Hexa8 grid with SVL's tag order (Builder.py:135-141 nodes, :167-183 elements).

    Node tag = (nx+1)(ny+1)k + (nx+1)j + i + 1; element connectivity
    n1..n8 = (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k), then the same at k+1.
    """
    P0, P1, P2, P3 = (np.asarray(p, dtype=np.float64) for p in (P0, P1, P2, P3))
    DX = (P1 - P0) / nx
    DY = (P2 - P0) / ny
    DZ = (P3 - P0) / nz
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = (P0[None, :] + i.reshape(-1, 1) * DX[None, :] + j.reshape(-1, 1) * DY[None, :]
              + k.reshape(-1, 1) * DZ[None, :])
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kk, jj, ii = kk.reshape(-1), jj.reshape(-1), ii.reshape(-1)
    sxy = (nx + 1) * (ny + 1)
    n1 = sxy * kk + (nx + 1) * jj + ii
    n4 = sxy * kk + (nx + 1) * (jj + 1) + ii
    conn = np.stack([n1, n1 + 1, n4 + 1, n4, n1 + sxy, n1 + 1 + sxy, n4 + 1 + sxy, n4 + sxy],
                    axis=1).astype(np.int32)
    # Find3DBoundaries (Builder.py:716-773).
    tolx, toly, tolz = DX[0] / 1000.0, DY[1] / 1000.0, DZ[2] / 1000.0
    top, bottom = P3[2] - tolz, P0[2] + tolz
    left, right = P0[0] + tolx, P1[0] - tolx
    front, back = P2[1] - toly, P0[1] + toly
    x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
    bnd = {
        "right": np.nonzero(x - right > 0.0)[0],
        "left": np.nonzero((x - right <= 0.0) & (left - x > 0.0))[0],
        "top": np.nonzero(z - top > 0.0)[0],
        "bottom": np.nonzero((z - top <= 0.0) & (bottom - z > 0.0))[0],
        "front": np.nonzero(y - front > 0.0)[0],
        "back": np.nonzero((y - front <= 0.0) & (back - y > 0.0))[0],
    }
    return coords, conn, bnd


def lysmer_dashpots(ndim: int, bnd: dict, names, eta_of):
    """Dashpot list equivalent to Builder.py:1086-1131 ``setLysmerDomain``.

    For every boundary node (first occurrence order over ``names``, ascending node
    inside each set) one restrained copy is made; for every boundary the node lies on,
    ``ndim`` ZeroLength1D dashpots are added, one per direction. ``eta_of(node, name,
    dir)`` returns the viscous coefficient; SVL's material index tables
    ``ind = [[0,1],[1,0]]`` (2D) and ``[[0,1,2],[2,0,1],[0,2,1]]`` (3D) mean the
    direction normal to the boundary takes the P-wave dashpot and the others the S-wave
    one. Corner and edge nodes are counted once per boundary, as in SVL.

    Returns (dash_node, dash_dir, dash_eta) in dashpot element-tag order.
    """
    order = []
    seen = {}
    for name in names:
        for n in np.sort(bnd[name]):
            n = int(n)
            if n not in seen:
                seen[n] = [name]
                order.append(n)
            else:
                seen[n].append(name)
    dn, dd, de = [], [], []
    for n in order:
        for name in seen[n]:
            for k in range(ndim):
                dn.append(n)
                dd.append(k)
                de.append(eta_of(n, name, k))
    return (np.asarray(dn, dtype=np.int32), np.asarray(dd, dtype=np.int32),
            np.asarray(de, dtype=np.float64))


def normal_dir(ndim: int, name: str) -> int:
    if ndim == 2:
        return 1 if name in ("top", "bottom") else 0
    return {"top": 2, "bottom": 2, "left": 0, "right": 0, "front": 1, "back": 1}[name]


def ricker_velocity(t, t0: float, f0: float, Ap: float):
    """Ricker pulse, velocity form: Ap (1 - 2 beta) exp(-beta), beta = (pi f0 (t - t0))^2.

    Closed form of 01-Pre_Process/Core/PlaneWave.py:222-223.
    """
    t = np.asarray(t, dtype=np.float64)
    beta = np.square(np.pi * f0 * (t - t0))
    return Ap * np.multiply(1.0 - 2.0 * beta, np.exp(-beta))


def tags_for(n: int) -> np.ndarray:
    return np.arange(1, n + 1, dtype=np.int64)
