"""Shared builders of the PML tests: the oracle's MixedModel for a PmlMesh."""
import numpy as np

from oracle import pml2d
from svl_amd.pml import PML2DQUAD4, PmlMesh


def to_oracle(m: PmlMesh) -> pml2d.MixedModel:
    elems = []
    for e in range(len(m.conn)):
        if m.elem_class[e] == PML2DQUAD4:
            p = (float(m.th[e]),) + tuple(float(v) for v in m.pml[e])
            elems.append(("pml", tuple(int(c) for c in m.conn[e]), m.E[e], m.nu[e], m.rho[e], p))
        else:
            elems.append(("quad", tuple(int(c) for c in m.conn[e]), m.E[e], m.nu[e], m.rho[e], (float(m.th[e]),)))
    return pml2d.MixedModel(np.asarray(m.coords), [int(k) for k in m.ndof], [list(f) for f in m.fixed], elems,
                            list(m.equal), lumped=bool(m.lumped), loads=list(m.loads))


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))
