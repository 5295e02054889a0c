"""Independent meshes advanced together on one GPU (no SVL counterpart: SVL runs one model per
process). The meshes are concatenated into one model whose parts share no node; the device
steps all parts with one fused launch per step (replayed from a captured hipGraph), each
part with its own time step and load series, and an autotuner picks the node-tile size.
"""
from __future__ import annotations

import numpy as np

from . import _capi as capi
from .integrators import DeviceModel
from .drm import attach
from .mesh import Model, tags_for


def union(models) -> tuple[Model, np.ndarray, np.ndarray]:
    """Concatenates models (same dimension, element type, thickness and mass form) into one
    model, part after part. Returns (model, part_node_ptr, part_free_ptr): the nodes and the
    free DOFs of part p are the ranges [ptr[p], ptr[p+1]) of the union (Plain numbering
    follows the node order, so each part's free DOFs stay contiguous and in their order)."""
    m0 = models[0]
    for m in models:
        if (m.ndim, m.elem_type, m.thickness, m.massform_lumped) != (m0.ndim, m0.elem_type, m0.thickness,
                                                                     m0.massform_lumped):
            raise ValueError("models of a batch must share dimension, element type, thickness and mass form")
    drm_parts = [m for m in models if m.drm_elem is not None and len(m.drm_elem)]
    if len({m.drm_factor for m in drm_parts}) > 1:
        raise ValueError("the domain reduction loads of a batch must share one factor")
    node_ptr = np.zeros(len(models) + 1, np.int32)
    free_ptr = np.zeros(len(models) + 1, np.int32)
    for i, m in enumerate(models):
        node_ptr[i + 1] = node_ptr[i] + m.nnodes
        free_ptr[i + 1] = free_ptr[i] + m.nfree
    nt = max(m.load_series.shape[1] if m.load_node.size else 1 for m in models)

    def cat(get):
        return np.concatenate([get(m) for m in models])

    def opt(get, m):
        v = get(m)
        return np.zeros(m.nelem) if v is None else np.asarray(v, dtype=np.float64)

    series = []
    for m in models:
        if m.load_node.size:
            s = np.zeros((m.load_node.size, nt))
            s[:, :m.load_series.shape[1]] = m.load_series
            series.append(s)
    u = Model(ndim=m0.ndim, coords=cat(lambda m: m.coords), node_tags=tags_for(int(node_ptr[-1])),
              elem_type=m0.elem_type,
              conn=np.concatenate([m.conn + node_ptr[i] for i, m in enumerate(models)]),
              E=cat(lambda m: m.E), nu=cat(lambda m: m.nu), rho=cat(lambda m: m.rho), thickness=m0.thickness,
              restrained=cat(lambda m: m.restrained), mat_type=cat(lambda m: m.mat_type), j2=cat(lambda m: m.j2),
              ray_am=np.concatenate([opt(lambda q: q.ray_am, m) for m in models]),
              ray_ak=np.concatenate([opt(lambda q: q.ray_ak, m) for m in models]),
              dash_node=np.concatenate([m.dash_node + node_ptr[i] for i, m in enumerate(models)]).astype(np.int32),
              dash_dir=cat(lambda m: m.dash_dir).astype(np.int32), dash_eta=cat(lambda m: m.dash_eta),
              load_node=np.concatenate([m.load_node + node_ptr[i] for i, m in enumerate(models)]).astype(np.int32),
              load_dir=(np.concatenate([m.load_dir.reshape(-1, m0.ndim) for m in models if m.load_node.size])
                        if any(m.load_node.size for m in models) else np.zeros((0, m0.ndim))),
              load_factor=cat(lambda m: m.load_factor),
              load_series=np.concatenate(series) if series else np.zeros((0, 1)),
              massform_lumped=m0.massform_lumped, name=f"batch_{len(models)}")
    if drm_parts:
        # One domain reduction load over every part's layer (parts share no node), element
        # and node indices offset per part, signals padded with zero rows.
        elem_off = np.concatenate([[0], np.cumsum([m.nelem for m in models])]).astype(np.int64)
        nts = max(m.drm_signal.shape[1] for m in drm_parts)
        el, nd, cd, sg = [], [], [], []
        for i, m in enumerate(models):
            if m.drm_elem is None or not len(m.drm_elem):
                continue
            el.append(np.asarray(m.drm_elem, np.int64) + elem_off[i])
            nd.append(np.asarray(m.drm_node, np.int64) + node_ptr[i])
            cd.append(np.asarray(m.drm_cond, np.int32))
            pad = np.zeros((m.drm_node.size, nts, m.drm_signal.shape[2]))
            pad[:, :m.drm_signal.shape[1]] = m.drm_signal
            sg.append(pad)
        attach(u, np.concatenate(el), np.concatenate(nd), np.concatenate(cd), np.concatenate(sg),
               drm_parts[0].drm_factor)
    return u, node_ptr, free_ptr


def autotune_tile_nodes(model: Model, part_node_ptr, dts, candidates=(128, 192, 256, 320, 448, 640, 896),
                        steps: int = 20, variant: int = 0):
    """Times `steps` fused CentralDifference steps for every candidate node-tile size on the
    device and returns (best, {candidate: microseconds per step or None})."""
    desc, keep = capi.make_desc(model)
    cand = np.ascontiguousarray(candidates, dtype=np.int32)
    ptr = np.ascontiguousarray(part_node_ptr, dtype=np.int32)
    dt = np.ascontiguousarray(dts, dtype=np.float64)
    us = np.zeros(cand.size)
    best = np.zeros(1, np.int32)
    import ctypes as C
    opts = capi.ExecOpts(0, 0, variant, 0, 0)
    capi.check(capi.lib().svlg_autotune_tiles(C.byref(desc), C.byref(opts), capi.iptr(cand), int(cand.size),
                                              int(dt.size), capi.iptr(ptr), capi.fptr(dt), int(steps),
                                              capi.iptr(best), capi.fptr(us)), "svlg_autotune_tiles")
    return int(best[0]), {int(c): (float(t) if t > 0 else None) for c, t in zip(cand, us)}


class BatchedCentralDifference:
    """CentralDifference over every part of a union model, each part with its own dt."""

    def __init__(self, mesh: DeviceModel, dts, part_node_ptr, part_free_ptr, ftol: float = 1e-12):
        self.mesh = mesh
        self.dts = np.ascontiguousarray(dts, dtype=np.float64)
        self.node_ptr = np.ascontiguousarray(part_node_ptr, dtype=np.int32)
        self.free_ptr = np.asarray(part_free_ptr)
        self.ftol = float(ftol)

    def Initialize(self):
        capi.check(capi.lib().svlg_cd_init_batch(self.mesh.h, int(self.dts.size), capi.iptr(self.node_ptr),
                                                 capi.fptr(self.dts), self.ftol), "svlg_cd_init_batch")

    def advance(self, k0: int, nsteps: int, graph_steps: int = 0, rec_dofs=None):
        """nsteps steps from step index k0: replayed hipGraphs of graph_steps launches (0: one
        eager launch per step; recording forces eager steps). Returns the history of the
        recorded union free DOFs when rec_dofs is given."""
        if rec_dofs is not None and len(rec_dofs):
            rec = np.ascontiguousarray(rec_dofs, dtype=np.int32)
            hist = np.zeros((nsteps, rec.size))
            capi.check(capi.lib().svlg_cd_steps(self.mesh.h, int(k0), int(nsteps), int(rec.size), capi.iptr(rec),
                                                capi.fptr(hist)), "svlg_cd_steps")
            return hist
        if graph_steps:
            capi.check(capi.lib().svlg_cd_graph_steps(self.mesh.h, int(k0), int(nsteps), int(graph_steps)),
                       "svlg_cd_graph_steps")
        else:
            capi.check(capi.lib().svlg_cd_steps_async(self.mesh.h, int(k0), int(nsteps)), "svlg_cd_steps_async")
        self.mesh.sync()
        return None

    def _state(self):
        n = self.mesh.nfree
        U, V, A = np.zeros(n), np.zeros(n), np.zeros(n)
        capi.check(capi.lib().svlg_cd_get_state(self.mesh.h, capi.fptr(U), capi.fptr(V), capi.fptr(A)),
                   "svlg_cd_get_state")
        return U, V, A

    def GetDisplacements(self, part: int | None = None) -> np.ndarray:
        U = self._state()[0]
        return U if part is None else U[self.free_ptr[part]:self.free_ptr[part + 1]]

    def GetVelocities(self, part: int | None = None) -> np.ndarray:
        V = self._state()[1]
        return V if part is None else V[self.free_ptr[part]:self.free_ptr[part + 1]]

    def GetAccelerations(self, part: int | None = None) -> np.ndarray:
        A = self._state()[2]
        return A if part is None else A[self.free_ptr[part]:self.free_ptr[part + 1]]
