"""Shared builders of the PML tests: the oracle's MixedModel for a PmlMesh."""
import numpy as np

from oracle import pml2d
from svl_amd.pml import LIN3DHEXA8, PML2DQUAD4, PML3DHEXA8, PmlMesh


def to_oracle(m: PmlMesh) -> pml2d.MixedModel:
    elems = []
    for e in range(len(m.conn)):
        conn = tuple(int(c) for c in m.conn[e])
        q = [float(v) for v in m.pml[e]]
        if m.elem_class[e] == PML2DQUAD4:
            p = (float(m.th[e]), q[0], q[1], q[2], q[3], q[4], q[6], q[7])
            elems.append(("pml", conn, m.E[e], m.nu[e], m.rho[e], p))
        elif m.elem_class[e] == PML3DHEXA8:
            elems.append(("pml3", conn, m.E[e], m.nu[e], m.rho[e], tuple(q)))
        elif m.elem_class[e] == LIN3DHEXA8:
            elems.append(("hexa", conn, m.E[e], m.nu[e], m.rho[e], ()))
        else:
            elems.append(("quad", tuple(int(c) for c in m.conn[e]), m.E[e], m.nu[e], m.rho[e], (float(m.th[e]),)))
    return pml2d.MixedModel(np.asarray(m.coords), [int(k) for k in m.ndof], [list(f) for f in m.fixed], elems,
                            list(m.equal), lumped=bool(m.lumped), loads=list(m.loads))


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))
