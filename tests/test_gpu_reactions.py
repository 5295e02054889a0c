"""GPU parity of the reactions (svlg_reactions) against the oracle's
Integrator::ComputeReactionForce restatement, after every step, at every fixed mesh node:
relative L2 over the history <= 1e-9 (implicit: CG at 1e-13), <= 1e-10 (explicit).
"""
import numpy as np
import pytest

import oracle
from svl_amd import workloads as W
from svl_amd.integrators import CentralDifference, DeviceModel, ExtendedNewmarkBeta, NewmarkBeta

from _meshes import block_model

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fixed(m):
    nodes = np.nonzero(m.restrained.any(1))[0]
    tot = [int(n) * m.ndim + c for n in nodes for c in range(m.ndim)]
    return nodes, tot


def device_history(dm, integ, nodes, nt):
    out = []
    for k in range(1, nt):
        integ.advance(k, 1)
        out.append(dm.reactions(nodes, k).reshape(-1))
    return np.asarray(out)


def test_cd_reactions():
    m, dt, nt = W.layer2d(14, 7, nt=120, lysmer=False)
    bottom = np.nonzero(np.isclose(m.coords[:, 1], m.coords[:, 1].min()))[0]
    m.restrained[bottom] = True
    m.ray_am = np.full(m.nelem, 0.8)
    m.add_point_load(int(bottom[3]), [1.0, -1.0], np.cos(np.arange(nt) * 0.1) * 1e4)
    nodes, tot = fixed(m)
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, [], reac_tot=tot)["R"]
    dm = DeviceModel(m)
    cd = CentralDifference(dm, dt)
    cd.Initialize()
    got = device_history(dm, cd, nodes, nt)
    dm.close()
    err = rel(got, ref)
    print(f"CentralDifference reactions: rel L2 {err:.2e}")
    assert err < 1e-10


def test_newmark_reactions():
    m, dt, nt, src = block_model(4, 3, 3)
    m.ray_am = np.full(m.nelem, 0.5)
    m.ray_ak = np.full(m.nelem, 1e-3)
    nodes, tot = fixed(m)
    m.add_point_load(int(nodes[4]), [1.0, 0.5, 0.0], np.sin(np.arange(400) * 0.05) * 3e5)
    nt = 40
    ref = oracle.run(m, oracle.EXTENDED_NEWMARK, dt, nt, [], reac_tot=tot)["R"]
    dm = DeviceModel(m)
    nm = ExtendedNewmarkBeta(dm, dt, rtol=1e-13, maxit=5000)
    nm.Initialize()
    got = device_history(dm, nm, nodes, nt)
    dm.close()
    err = rel(got, ref)
    print(f"ExtendedNewmarkBeta reactions: rel L2 {err:.2e}")
    assert err < 1e-9


def test_newton_j2_reactions():
    m, dt, nt = W.soil_structure_j2(8, 8, 4, nt=31, block=(2, 2, 2), Ap=2e7, t0=0.05, f0=8.0)
    nodes, tot = fixed(m)
    ref = oracle.run(m, oracle.NEWMARK, dt, nt, [], algorithm=oracle.NEWTON, nmax=20, cnvgtest=5, cnvgtol=1e-6,
                     reac_tot=tot)
    dm = DeviceModel(m)
    nl = NewmarkBeta(dm, dt, cnvgtol=1e-6, nmax=20, cnvgtest=5, rtol=1e-13, maxit=5000)
    nl.Initialize()
    got = []
    its = []
    for k in range(1, nt):
        nl.advance(k, 1)
        its.append(int(nl.iters[0]))
        got.append(dm.reactions(nodes, k).reshape(-1))
    dm.close()
    assert np.array_equal(its, ref["iters"])
    assert ref["iters"].max() > 2
    err = rel(np.asarray(got), ref["R"])
    print(f"Newton J2 reactions: rel L2 {err:.2e}")
    assert err < 1e-9
