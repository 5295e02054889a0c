"""CPU checks of the node-tile plan used by the fused explicit kernel.

svlg_plan_tiles builds the tiles on the host and verifies, inside the library, that
every node is owned by exactly one tile (plain nodes first), that each tile lists
exactly the elements touching its nodes in color order, and that every element corner
maps to a staged slot holding that corner's node. These tests run it on structured,
rotated, distorted and randomly renumbered meshes and check the returned sizes.
"""
import numpy as np
import pytest

from svl_amd import workloads as W
from svl_amd.integrators import plan_tiles
from svl_amd.mesh import ELEM_HEXA8, Model, make_domain_volume, tags_for


def renumbered(m: Model, seed: int) -> Model:
    """Same mesh with node indices shuffled (an unstructured-looking numbering)."""
    rng = np.random.default_rng(seed)
    p = rng.permutation(m.nnodes)          # new -> old
    inv = np.empty_like(p)
    inv[p] = np.arange(m.nnodes)
    return Model(ndim=m.ndim, coords=m.coords[p], node_tags=tags_for(m.nnodes), elem_type=m.elem_type,
                 conn=inv[m.conn], E=m.E, nu=m.nu, rho=m.rho, thickness=m.thickness,
                 restrained=m.restrained[p], dash_node=inv[m.dash_node], dash_dir=m.dash_dir,
                 dash_eta=m.dash_eta)


@pytest.mark.parametrize("tile_nodes", [32, 100, 448, 2000])
def test_plan_structured_3d(tile_nodes):
    m, _, _ = W.halfspace3d(12, 10, 7, h=2.0, nt=4)
    st = plan_tiles(m, tile_nodes)
    assert st["max_own"] <= tile_nodes
    assert st["ncolors"] == 8
    assert st["tile_elems"] >= m.nelem
    assert st["ntiles"] >= int(np.ceil(m.nnodes / tile_nodes))


def test_plan_2d_and_renumbered():
    m, _, _ = W.layer2d(30, 17, nt=4)
    st = plan_tiles(m, 64)
    assert st["ncolors"] == 4 and st["max_own"] <= 64
    r = renumbered(m, 3)
    st2 = plan_tiles(r, 64)
    assert st2["max_own"] <= 64 and st2["tile_elems"] >= r.nelem


def test_plan_distorted_and_restrained():
    coords, conn, bnd = make_domain_volume(6, 5, 4, (0, 0, 0), (6, 0, 0), (0, 5, 0), (0, 0, 4))
    rng = np.random.default_rng(1)
    interior = np.all((coords > 0.01) & (coords < np.array([5.99, 4.99, 3.99])), axis=1)
    coords = coords + interior[:, None] * rng.uniform(-0.2, 0.2, coords.shape)
    restrained = np.zeros(coords.shape, bool)
    restrained[bnd["bottom"]] = True
    m = Model(ndim=3, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_HEXA8, conn=conn,
              E=1e8, nu=0.3, rho=2000.0, restrained=restrained)
    st = plan_tiles(m, 40)
    assert st["max_own"] <= 40


def test_compact_tiles_on_structured_grid():
    """Plane-aligned bisection keeps the halo of a ~7^3-node tile near 9^3 - 7^3."""
    m, _, _ = W.halfspace3d(40, 40, 20, h=2.0, nt=4)
    st = plan_tiles(m, 448)
    avg_stage = (m.nnodes + st["halo_nodes"]) / st["ntiles"]
    avg_own = m.nnodes / st["ntiles"]
    assert avg_stage < 2.3 * avg_own
    assert st["redundancy"] < 1.6
