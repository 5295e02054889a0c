"""GPU parity of the fused node-tile CentralDifference kernel.

Covers the affine modal path (axis-aligned, rotated + sheared parallelepipeds) and the
Gauss-point path (distorted elements), several tile sizes (so tiles with halos, a
single tile, tiny tiles), restrained nodes, dashpots and loads. Bar: relative L2 error
of the recorded displacement history and of the final field <= 1e-10 against the FP64
oracle; the fused kernel agrees with the element-colored kernels to 1e-12.
"""
import numpy as np
import pytest

import oracle
from svl_amd import workloads as W
from svl_amd.integrators import CentralDifference, DeviceModel
from svl_amd.mesh import ELEM_HEXA8, ELEM_QUAD4, Model, make_domain_area, make_domain_volume, ricker_velocity, tags_for

from _meshes import block_model, graded, rotation

pytestmark = pytest.mark.gpu

REL_TOL = 1e-10


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_gpu(m, dt, nsteps, rec, **opts):
    dm = DeviceModel(m, **opts)
    info = dm.info()
    cd = CentralDifference(dm, dt)
    cd.Initialize()
    hist = cd.advance(1, nsteps, rec)
    U = cd.GetDisplacements()
    dm.close()
    return hist, U, info


@pytest.mark.parametrize("kind", ["axis", "rotated", "graded", "distorted"])
@pytest.mark.parametrize("tile_nodes,geo", [(24, 0), (120, 0), (120, -1), (300, 0)])
def test_fused_hexa_parity(kind, tile_nodes, geo):
    if kind == "axis":
        m, dt, nt, src = block_model(7, 6, 5)
    elif kind == "rotated":
        m, dt, nt, src = block_model(7, 6, 5, transform=rotation())
    elif kind == "graded":
        m, dt, nt, src = block_model(7, 6, 5)
        m = graded(m)
    else:
        m, dt, nt, src = block_model(7, 6, 5, distort=0.18, seed=4)
    rec = [d for d in range(0, m.nfree, 11)]
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    hist, U, info = run_gpu(m, dt, nt - 1, rec, tile_nodes=tile_nodes, geo_table=geo)
    assert info["variant"] == 3
    expect = 0 if kind == "distorted" else (1 if geo == -1 else (3 if kind in ("axis", "graded") else 2))
    assert info["elem_path"] == expect
    assert rel_l2(hist, ref["U"]) < REL_TOL
    assert rel_l2(U, ref["final_U"]) < REL_TOL


@pytest.mark.parametrize("tile_nodes", [16, 200, 4000])
def test_fused_quad_layer_parity(tile_nodes):
    m, dt, nt = W.layer2d(24, 12, nt=1201)
    rec = W.dofs_of(m, int(m.load_node[0])) + list(range(3, m.nfree, 17))
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    for geo in (0, -1):
        hist, U, info = run_gpu(m, dt, nt - 1, rec, tile_nodes=tile_nodes, geo_table=geo)
        assert info["variant"] in (1, 3) and info["elem_path"] == (3 if geo == 0 else 1)
        assert rel_l2(hist, ref["U"]) < REL_TOL


def test_fused_quad_distorted():
    coords, conn, bnd = make_domain_area(10, 6, (0, 0), (10, 0), (0, 6))
    rng = np.random.default_rng(9)
    inner = ~np.isin(np.arange(coords.shape[0]), np.concatenate(list(bnd.values())))
    coords[inner] += rng.uniform(-0.2, 0.2, (int(inner.sum()), 2))
    restrained = np.zeros(coords.shape, bool)
    restrained[bnd["bottom"]] = True
    m = Model(ndim=2, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_QUAD4, conn=conn,
              E=1e8, nu=0.25, rho=2000.0, thickness=0.7, restrained=restrained)
    nt, dt = 800, 2e-4
    m.add_point_load(int(bnd["top"][3]), [0.0, -1.0], ricker_velocity(np.arange(nt) * dt, 0.03, 20.0, 1e5))
    rec = list(range(0, m.nfree, 5))
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    hist, U, info = run_gpu(m, dt, nt - 1, rec, tile_nodes=20)
    assert info["affine"] == 0
    assert rel_l2(hist, ref["U"]) < REL_TOL


@pytest.mark.parametrize("variant", [1, 3])
def test_fused_matches_colored_kernels(variant):
    """Both fused flavours (colour phases, corner slots) against the element-colored kernels."""
    m, dt, _ = W.halfspace3d(14, 12, 8, h=2.0, nt=301)
    rec = list(range(0, m.nfree, 13))
    h1, U1, i1 = run_gpu(m, dt, 300, rec, tile_nodes=128, variant=variant)
    h2, U2, i2 = run_gpu(m, dt, 300, rec, variant=2)
    assert i1["variant"] == variant and i2["variant"] == 2
    assert rel_l2(U1, U2) < 1e-12
    assert rel_l2(h1, h2) < 1e-12


def test_colour_phase_kernel_on_non_corner_unique_mesh():
    """Two blocks glued with mirrored node order break corner uniqueness: auto picks the
    colour-phase kernel, which still matches the oracle."""
    coords, conn, bnd = make_domain_volume(6, 5, 4, (0, 0, 0), (6, 0, 0), (0, 5, 0), (0, 0, 4))
    conn = conn.copy()
    # Re-orient every other element by a rotation of its corners about z (still a valid
    # right-handed hexahedron), so a node can be corner 0 of two elements.
    rot = np.array([1, 2, 3, 0, 5, 6, 7, 4])
    conn[::2] = conn[::2][:, rot]
    restrained = np.zeros(coords.shape, bool)
    restrained[bnd["bottom"]] = True
    m = Model(ndim=3, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_HEXA8, conn=conn,
              E=1e8, nu=0.3, rho=2000.0, restrained=restrained)
    nt, dt = 300, 2e-4
    m.add_point_load(int(bnd["top"][5]), [0.2, 0.1, -1.0], ricker_velocity(np.arange(nt) * dt, 0.02, 25.0, 1e6))
    rec = list(range(0, m.nfree, 7))
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    hist, U, info = run_gpu(m, dt, nt - 1, rec, tile_nodes=40)
    assert info["variant"] == 1
    assert rel_l2(hist, ref["U"]) < REL_TOL
    with pytest.raises(Exception):
        run_gpu(m, dt, 2, rec, tile_nodes=40, variant=3)


def test_fused_deterministic():
    m, dt, _ = W.halfspace3d(10, 9, 6, h=2.0, nt=201)
    _, U1, _ = run_gpu(m, dt, 200, [0], tile_nodes=100)
    _, U2, _ = run_gpu(m, dt, 200, [0], tile_nodes=100)
    assert np.array_equal(U1, U2)


def test_bf16_mfma_element_variant():
    """BF16 MFMA element step (FP32 accumulation, FP64 state and update): the displacement
    history stays within 1e-2 relative L2 of the FP64 oracle (measured ~1e-3) and the
    variant is refused where its single-geometry condition does not hold."""
    m, dt, nt = W.halfspace3d(12, 10, 8, h=2.0, nt=401)
    rec = W.dofs_of(m, int(m.load_node[0])) + list(range(0, m.nfree, 13))
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    hist, U, info = run_gpu(m, dt, nt - 1, rec, tile_nodes=120, elem_precision=1)
    assert info["elem_precision"] == 1 and info["ngeo"] == 1
    err = rel_l2(hist, ref["U"])
    print(f"bf16 element step: history rel L2 error vs FP64 oracle {err:.3e}")
    assert err < 1e-2
    h64, U64, _ = run_gpu(m, dt, nt - 1, rec, tile_nodes=120)
    assert rel_l2(h64, ref["U"]) < REL_TOL
    mg = graded(block_model(5, 4, 4)[0])
    with pytest.raises(Exception):
        DeviceModel(mg, elem_precision=1)


@pytest.mark.parametrize("tile_nodes", [120, 0])
def test_fp32_element_variant(tile_nodes):
    """FP32 element step (modal force in single precision from the FP64 staged displacements;
    FP64 corner-slot sums, state and update) on the layered half-space over 400 steps: the
    displacement history stays within 1e-4 relative L2 of the FP64 oracle, graded bricks run it,
    and meshes that are not axis-aligned bricks are refused."""
    m, dt, nt = W.halfspace3d(12, 10, 8, h=2.0, nt=401)
    rec = W.dofs_of(m, int(m.load_node[0])) + list(range(0, m.nfree, 13))
    ref = oracle.run(m, oracle.CENTRAL_DIFFERENCE, dt, nt, rec)
    hist, U, info = run_gpu(m, dt, nt - 1, rec, tile_nodes=tile_nodes, elem_precision=2)
    assert info["elem_precision"] == 2
    err = rel_l2(hist, ref["U"])
    print(f"fp32 element step (tiles {tile_nodes}): history rel L2 error vs FP64 oracle {err:.3e}")
    assert err < 1e-4
    mb, dtg, _, _ = block_model(5, 4, 4)
    mg = graded(mb)
    refg = oracle.run(mg, oracle.CENTRAL_DIFFERENCE, dtg, 201, list(range(0, mg.nfree, 7)))
    hg, _, ig = run_gpu(mg, dtg, 200, list(range(0, mg.nfree, 7)), tile_nodes=tile_nodes, elem_precision=2)
    assert ig["elem_precision"] == 2 and rel_l2(hg, refg["U"]) < 1e-4
    with pytest.raises(Exception):
        DeviceModel(block_model(4, 4, 3, transform=rotation())[0], elem_precision=2)
