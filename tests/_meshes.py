"""Small meshes shared by the GPU parity tests (no GPU needed to build them)."""
import numpy as np

from svl_amd.mesh import ELEM_HEXA8, ELEM_QUAD4, Model, make_domain_area, make_domain_volume, ricker_velocity, tags_for


def block_model(nx, ny, nz, transform=None, distort=0.0, seed=0):
    """Hexa8 block, base restrained, dashpots on the sides, Ricker load on top."""
    coords, conn, bnd = make_domain_volume(nx, ny, nz, (0, 0, 0), (nx, 0, 0), (0, ny, 0), (0, 0, nz))
    coords = coords.astype(np.float64)
    if distort:
        rng = np.random.default_rng(seed)
        inner = ~np.isin(np.arange(coords.shape[0]), np.concatenate(list(bnd.values())))
        coords[inner] += rng.uniform(-distort, distort, (int(inner.sum()), 3))
    if transform is not None:
        coords = coords @ transform.T
    restrained = np.zeros(coords.shape, bool)
    restrained[bnd["bottom"]] = True
    rng = np.random.default_rng(seed + 1)
    E = rng.uniform(5e7, 2e8, conn.shape[0])
    m = Model(ndim=3, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_HEXA8, conn=conn,
              E=E, nu=0.3, rho=2000.0, restrained=restrained)
    side = np.unique(np.concatenate([bnd["left"], bnd["right"], bnd["front"], bnd["back"]]))
    side = side[~np.isin(side, bnd["bottom"])]
    m.dash_node = np.repeat(side, 3).astype(np.int32)
    m.dash_dir = np.tile(np.arange(3), side.size).astype(np.int32)
    m.dash_eta = np.full(side.size * 3, 3.0e5)
    nt = 400
    dt = 2e-4
    top = bnd["top"]
    src = int(top[np.argmin(np.linalg.norm(coords[top] - coords[top].mean(0), axis=1))])
    m.add_point_load(src, [0.3, 0.2, -1.0], ricker_velocity(np.arange(nt) * dt, 0.02, 25.0, 1e6))
    m.add_point_load(int(top[0]), [1.0, 0.0, 0.0], ricker_velocity(np.arange(nt) * dt, 0.03, 20.0, 5e5), 0.5)
    return m, dt, nt, src


def rotation():
    a, b = 0.3, -0.45
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    S = np.array([[1.0, 0.2, 0.0], [0.0, 1.1, 0.15], [0.0, 0.0, 0.9]])
    return Rz @ Rx @ S


def graded(m):
    """Same block with a non-uniform spacing along every axis (still parallelepipeds,
    one geometry per element column/row/layer combination)."""
    c = m.coords.copy()
    for d in range(3):
        vals = np.unique(c[:, d])
        new = np.cumsum(np.concatenate([[0.0], 0.7 + 0.6 * np.random.default_rng(d).random(vals.size - 1)]))
        c[:, d] = new[np.searchsorted(vals, c[:, d])]
    m.coords = np.ascontiguousarray(c)
    return m


def drm_quad_model(nt, dt, half=(3.5, 2.5)):
    """Quad4 grid with every side fixed, a source outside a DRM box (center, half sides)."""
    coords, conn, bnd = make_domain_area(16, 12, (0, 0), (16, 0), (0, 12))
    restrained = np.zeros(coords.shape, bool)
    for k in ("left", "right", "top", "bottom"):
        restrained[bnd[k]] = True
    rng = np.random.default_rng(3)
    m = Model(ndim=2, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_QUAD4, conn=conn,
              E=rng.uniform(0.8e8, 1.2e8, conn.shape[0]), nu=0.3, rho=2000.0, restrained=restrained)
    src = int(np.argmin(np.linalg.norm(coords - np.array([2.0, 10.0]), axis=1)))
    m.add_point_load(src, [1.0, -0.5], ricker_velocity(np.arange(nt) * dt, 0.03, 30.0, 1e6))
    return m, ((8.0, 6.0), half)


def drm_hexa_model(nt, dt, n=(8, 8, 6)):
    """Hexa8 block with every side fixed, a source outside a DRM box (center, half sides)."""
    coords, conn, bnd = make_domain_volume(*n, (0, 0, 0), (n[0], 0, 0), (0, n[1], 0), (0, 0, n[2]))
    restrained = np.zeros(coords.shape, bool)
    for k in bnd:
        restrained[bnd[k]] = True
    m = Model(ndim=3, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_HEXA8, conn=conn,
              E=1e8, nu=0.25, rho=2000.0, restrained=restrained)
    src = int(np.argmin(np.linalg.norm(coords - np.array([1.0, 1.0, n[2] - 1.0]), axis=1)))
    m.add_point_load(src, [0.3, 0.5, -1.0], ricker_velocity(np.arange(nt) * dt, 0.02, 40.0, 1e6))
    c = np.array(n, float) / 2
    return m, ((c[0], c[1] + 0.5, c[2]), (n[0] / 5.3, n[1] / 5.3, n[2] / 4))
