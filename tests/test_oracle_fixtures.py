"""The FP64 oracle against SVL's own validation cases (CPU).

SVL's debugging cases F02, J02, F06 and J05 compare SVL against OpenSees runs of the
same model; the OpenSees histories are stored with 6 significant digits (the fixtures
in tests/golden/, written by tests/golden/make_fixtures.py). The models are rebuilt
here from the case definitions (node/element tags, restraints, loads, damping), run
through the oracle's Newmark (gamma=1/2, beta=1/4, Linear algorithm, factor-once
direct solve) with SVL's index convention (row k-1 holds step k), and compared over
the whole history. The measured agreement is 1-2e-6 relative RMS, the floor set by the
6-digit reference; the tests assert 1e-5. This pins the element stiffness, the lumped
and consistent mass, Rayleigh damping, dashpots, loads and the integrator, not to 1e-10:
the 1e-10 bar applies between the oracle and the device path.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from svl_amd.mesh import ELEM_HEXA8, ELEM_QUAD4, Model, make_domain_volume, tags_for

GOLD = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLD / f"{name}.npz") as z:
        return {k: z[k] for k in z.files}


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def test_F02_quad4_point_load_newmark():
    g = load("F02")
    coords = np.array([[0.0, 0.0], [2.0, 0.0], [0.0, 2.0], [2.0, 2.0]])
    restrained = np.array([[True, True], [False, True], [False, False], [False, False]])
    m = Model(ndim=2, coords=coords, node_tags=tags_for(4), elem_type=ELEM_QUAD4, conn=np.array([[0, 1, 3, 2]]),
              E=2.08e8, nu=0.30, rho=2000.0, thickness=1.0, restrained=restrained)
    m.add_point_load(3, [0.01, -0.02], g["load"])
    rec = [int(d) for d in m.free_dof()[3]]
    r = oracle.run(m, oracle.NEWMARK, 0.025, 201, rec, want_va=True)
    os_u = g["displacement"][:, 1:3]
    assert r["U"].shape == os_u.shape
    assert rel_rms(r["U"], os_u) < 1e-5
    assert rel_rms(r["V"], g["velocity"][:, 1:3]) < 1e-5
    assert rel_rms(r["A"], g["acceleration"][:, 1:3]) < 1e-5
    # Both Newmark forms of SVL agree for the Linear algorithm.
    r2 = oracle.run(m, oracle.EXTENDED_NEWMARK, 0.025, 201, rec)
    assert np.array_equal(r["U"], r2["U"])


def test_J02_hexa8_point_load_newmark():
    g = load("J02")
    coords = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2]],
                      dtype=float)
    restrained = np.zeros((8, 3), bool)
    restrained[0] = True
    restrained[1:4, 2] = True
    # J02 leaves massform at SVL's default, 'Consistent' (01-Pre_Process/Core/Definitions.py:45).
    m = Model(ndim=3, coords=coords, node_tags=tags_for(8), elem_type=ELEM_HEXA8,
              conn=np.arange(8, dtype=np.int32)[None, :], E=2.08e8, nu=0.30, rho=2000.0, restrained=restrained,
              massform_lumped=False)
    m.add_point_load(5, [0.01, 0.0, 0.0], g["load"])
    m.add_point_load(7, [0.0, 0.01, 0.0], g["load"])
    m.add_point_load(6, [0.0, 0.0, -0.05], g["load"])
    rec = [int(d) for d in m.free_dof()[6]]
    r = oracle.run(m, oracle.NEWMARK, 0.025, 201, rec)
    # The OpenSees model is written in a permuted frame: SVL (x, y, z) = OpenSees (3, 1, 2).
    os_u = g["displacement"][:, [3, 1, 2]]
    assert rel_rms(r["U"], os_u) < 1e-5
    # A plausible slip -- lumping the mass -- is visible at this resolution.
    m.massform_lumped = True
    assert rel_rms(oracle.run(m, oracle.NEWMARK, 0.025, 201, rec)["U"], os_u) > 1e-4


def soil_column_2d(g):
    n = 101
    coords = np.zeros((2 * n, 2))
    coords[0::2, 0], coords[1::2, 0] = 0.0, 1.0
    coords[0::2, 1] = coords[1::2, 1] = np.arange(n, dtype=float)
    k = np.arange(n - 1)
    conn = np.stack([2 * k, 2 * k + 1, 2 * k + 3, 2 * k + 2], axis=1).astype(np.int32)
    restrained = np.zeros(coords.shape, bool)
    restrained[:, 1] = True
    m = Model(ndim=2, coords=coords, node_tags=tags_for(2 * n), elem_type=ELEM_QUAD4, conn=conn, E=1.3e7, nu=0.30,
              rho=2000.0, thickness=1.0, restrained=restrained, ray_am=0.1244195110, ray_ak=0.0007878958,
              dash_node=np.array([0, 1]), dash_dir=np.array([0, 0]), dash_eta=np.array([5.0e4, 5.0e4]))
    m.add_point_load(0, [5.0e4, 0.0], g["load"])
    m.add_point_load(1, [5.0e4, 0.0], g["load"])
    return m


def test_F06_quad4_soil_column_rayleigh_dashpots():
    g = load("F06")
    m = soil_column_2d(g)
    top = 200
    rec = [int(m.free_dof()[top][0])]
    r = oracle.run(m, oracle.NEWMARK, 0.01, 1000, rec)
    assert rel_rms(r["U"][:, 0], g["displacement"][:, 3]) < 1e-5


def test_J05_hexa8_soil_column_rayleigh_dashpots():
    g = load("J05")
    coords, conn, bnd = make_domain_volume(1, 1, 100, (0, 0, 0), (1.0, 0, 0), (0, 1.0, 0), (0, 0, 100.0))
    restrained = np.zeros(coords.shape, bool)
    restrained[:, 2] = True
    dn = np.repeat(np.arange(4), 2)
    dd = np.tile([0, 1], 4)
    m = Model(ndim=3, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_HEXA8, conn=conn,
              E=1.3e7, nu=0.30, rho=2000.0, restrained=restrained, ray_am=0.1244195110, ray_ak=0.0007878958,
              dash_node=dn, dash_dir=dd, dash_eta=np.full(8, 2.5e4))
    for n in range(4):
        m.add_point_load(n, [25000.0, 0.0, 0.0], g["load"])
    top = 400  # node tag 401
    rec = [int(m.free_dof()[top][0])]
    r = oracle.run(m, oracle.NEWMARK, 0.01, 1000, rec)
    # The reference for this 3D column is SVL's 2D plane-strain OpenSees column (a 1D
    # shear column either way).
    assert rel_rms(r["U"][:, 0], g["displacement"][:, 3]) < 1e-5
