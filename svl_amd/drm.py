"""Domain reduction (SVL's ElementLoad GENERALWAVE) set-up on the host.

* :func:`drm_box` selects the DRM layer the way SVL's pre-processor does
  (01-Pre_Process/Method/Builder.py:1010-1073): an element is a DRM element when some but
  not all of its nodes lie in the closed box |x - x0| <= xl; its nodes inside the box are
  the interior (boundary) nodes, the others the exterior nodes.
* :func:`signals_from_history` turns the free-DOF histories of a background run on the same
  mesh into node signals with the integrator's row alignment, so that the reduced model
  reproduces that run inside the box and stays at rest outside it.
* :func:`attach` stores the load on a :class:`~svl_amd.mesh.Model`.

The device computes the forces (lin*::ComputeDomainReductionForces) every step; nothing
here evaluates them.
"""
from __future__ import annotations

import numpy as np

from .mesh import Model


def drm_box(m: Model, x0, xl):
    """Returns (elements, nodes, cond): DRM element indices in ascending order, the nodes of
    those elements in ascending order and their exterior flags (1 outside the box)."""
    x0 = np.asarray(x0, dtype=np.float64)
    xl = np.asarray(xl, dtype=np.float64)
    inside = np.all(np.abs(m.coords - x0[None, :]) <= xl[None, :], axis=1)
    cnt = inside[m.conn].sum(axis=1)
    elems = np.nonzero((cnt > 0) & (cnt < m.nen))[0].astype(np.int32)
    nodes = np.unique(m.conn[elems].reshape(-1)).astype(np.int32)
    cond = (~inside[nodes]).astype(np.int32)
    return elems, nodes, cond


def signals_from_history(m: Model, nodes, nt: int, U, V=None, A=None, lag: int = 1, cols=None):
    """Node signals [len(nodes), nt, 3*ndim] from free-DOF histories [rows, ncols] of a run
    whose row r holds the state after step r + 1 (the recorder convention); column i holds
    free DOF cols[i] (default: every free DOF in order). Row k of the signal is history row
    k - lag, zero before: lag = 1 for (Extended)NewmarkBeta, whose step k balances forces at
    the new state, lag = 2 for CentralDifference, whose step k evaluates the stiffness at the
    state of step k - 1. Restrained DOFs read zero."""
    nd = m.ndim
    fd = m.free_dof()
    nodes = np.asarray(nodes, dtype=np.int64)
    if cols is not None:
        pos = np.full(m.nfree, -1, np.int64)
        pos[np.asarray(cols, dtype=np.int64)] = np.arange(len(cols))
        need = fd[nodes]
        if np.any((need >= 0) & (pos[np.maximum(need, 0)] < 0)):
            raise ValueError("every free DOF of the signal nodes needs a history column")
        fd = np.where(fd >= 0, pos[np.maximum(fd, 0)], -1)
    sig = np.zeros((nodes.size, nt, 3 * nd))
    for f, H in enumerate((U, V, A)):
        if H is None:
            continue
        H = np.asarray(H, dtype=np.float64)
        rows = np.arange(nt) - lag
        ok = (rows >= 0) & (rows < H.shape[0])
        for c in range(nd):
            idx = fd[nodes, c]
            free = idx >= 0
            blk = np.zeros((nodes.size, nt))
            blk[np.ix_(free, ok)] = H[np.ix_(rows[ok], idx[free])].T
            sig[:, :, f * nd + c] = blk
    return sig


def attach(m: Model, elems, nodes, cond, signal, factor: float = 1.0) -> Model:
    """Sets the DRM load of ``m`` (signals as written in SVL's node files: the library
    negates the exterior ones, as SVL's reader does)."""
    elems = np.ascontiguousarray(elems, dtype=np.int32)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    cond = np.ascontiguousarray(cond, dtype=np.int32)
    signal = np.ascontiguousarray(signal, dtype=np.float64)
    if signal.ndim != 3 or signal.shape[0] != nodes.size or signal.shape[2] != 3 * m.ndim:
        raise ValueError("DRM signal must be [n_drm_nodes, nt, 3*ndim]")
    if cond.size != nodes.size:
        raise ValueError("one exterior flag per DRM node")
    need = np.unique(m.conn[elems].reshape(-1))
    if not np.all(np.isin(need, nodes)):
        raise ValueError("every node of a DRM element needs a signal")
    m.drm_elem, m.drm_node, m.drm_cond, m.drm_signal = elems, nodes, cond, signal
    m.drm_factor = float(factor)
    return m
