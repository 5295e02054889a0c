"""CPU checks of the 2D PML host side: SVL's EQUAL-aware free numbering against the oracle's
T matrix, the mesh builders, and the argument checks of svlg_pml_create (which run before any
device call)."""
import ctypes as C

import numpy as np
import pytest

from _pml_meshes import to_oracle
from svl_amd import _capi as capi
from svl_amd.pml import PmlMesh, PmlNewmark, halfspace_box, halfspace_box3d, soil_column, soil_rod3d


@pytest.mark.parametrize("build", [lambda: soil_column(90, 10, np.ones(3)), lambda: halfspace_box(6, 4, 2),
                                   lambda: soil_rod3d(20, 5, np.ones(3)), lambda: halfspace_box3d(3, 2, 2, 1)])
def test_numbering_matches_oracle_T(build):
    m = build()
    ptr, free, fac, nfree = m.numbering()
    off, T = to_oracle(m).numbering()
    assert np.array_equal(ptr, off) and T.shape[1] == nfree
    Tp = np.zeros_like(T)
    for d in range(T.shape[0]):
        if free[d] >= 0:
            Tp[d, free[d]] = fac[d]
    assert np.array_equal(Tp, T)


def test_halfspace_box_layout():
    m = halfspace_box(8, 5, 3)
    ne = (8 + 6) * (5 + 3)
    assert len(m.conn) == ne and (m.elem_class == 0).sum() == 8 * 5
    # every regular node on the PML interface is the master of two EQUAL constraints
    iface = 2 * (5 + 1) + (8 + 1) - 2
    assert len(m.equal) == 2 * iface
    for s, a, mm, b, f in m.equal:
        assert np.array_equal(m.coords[s], m.coords[mm]) and m.ndof[s] == 5 and m.ndof[mm] == 2 and a == b


def test_halfspace_box3d_layout():
    m = halfspace_box3d(4, 3, 2, 2)
    assert len(m.conn) == 8 * 7 * 4 and (m.elem_class == 2).sum() == 4 * 3 * 2
    # corner elements of the bottom layer are stretched in three directions
    three = [e for e in range(len(m.conn)) if m.elem_class[e] == 3 and np.all(m.pml[e, 6:] != 0)]
    assert len(three) == 4 * 2 * 2 * 2
    for s_, a, mm, b, f in m.equal:
        assert np.array_equal(m.coords[s_], m.coords[mm]) and m.ndof[s_] == 9 and m.ndof[mm] == 3


def _bad(mesh_edit):
    m = soil_column(4, 2, np.ones(3))
    mesh_edit(m)
    with pytest.raises(capi.SvlgError) as ei:
        PmlNewmark(m, 0.01)
    return str(ei.value)


@pytest.mark.parametrize("edit,msg", [
    (lambda m: m.pml.__setitem__((0, 2), 2.0), "0 < R <= 1"),
    (lambda m: m.pml.__setitem__((0, 1), 0.0), "PML depth"),
    (lambda m: m.elem_class.__setitem__(0, 0), "unknown count"),
    (lambda m: m.conn.__setitem__((0, 0), 999), "out of range"),
    (lambda m: m.nu.__setitem__(0, 0.5), "material"),
    (lambda m: m.elem_class.__setitem__(0, 3), "dimension"),
])
def test_create_rejects_bad_models(edit, msg):
    assert msg in _bad(edit)
