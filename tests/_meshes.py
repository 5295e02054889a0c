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


def base_motion_series(nt, dt, length=None, g0=5e-4, amp=2e-3, t1=0.04):
    """A base displacement history that starts at g0 (nonzero, so the sample-0 rule matters),
    rises by amp with a (1 - cos) pulse over t1 and is cut at `length` samples (past the end
    SVL reads the first value again)."""
    n = nt if length is None else length
    t = np.arange(n) * dt
    return g0 + amp * np.where(t < t1, 0.5 * (1.0 - np.cos(2.0 * np.pi * t / t1)), 0.0) + 1e-3 * np.minimum(t / t1, 1.0)


def support_model(ndim, nt, dt, am=0.0, ak=0.0, lumped=True, dashpots=False, load=True, length=None, factor=1.0,
                  dirs=(0,)):
    """Quad4 (10 x 6) or Hexa8 (5 x 4 x 4) block on a base whose nodes are restrained in every
    direction and move in `dirs` with base_motion_series (one SupportMotion load with the given
    combination factor); optional Rayleigh damping, consistent mass, Lysmer dashpots on the sides
    and a point load on top."""
    if ndim == 2:
        coords, conn, bnd = make_domain_area(10, 6, (0, 0), (10, 0), (0, 6))
        etype, side_names = ELEM_QUAD4, ("left", "right")
    else:
        coords, conn, bnd = make_domain_volume(5, 4, 4, (0, 0, 0), (5, 0, 0), (0, 4, 0), (0, 0, 4))
        etype, side_names = ELEM_HEXA8, ("left", "right", "front", "back")
    restrained = np.zeros(coords.shape, bool)
    restrained[bnd["bottom"]] = True
    rng = np.random.default_rng(11)
    m = Model(ndim=ndim, coords=coords.astype(np.float64), node_tags=tags_for(coords.shape[0]), elem_type=etype,
              conn=conn, E=rng.uniform(6e7, 1.4e8, conn.shape[0]), nu=0.3, rho=2000.0, restrained=restrained,
              massform_lumped=lumped)
    m.ray_am = np.full(conn.shape[0], am)
    m.ray_ak = np.full(conn.shape[0], ak)
    if dashpots:
        side = np.unique(np.concatenate([bnd[k] for k in side_names]))
        side = side[~np.isin(side, bnd["bottom"])]
        m.dash_node = np.repeat(side, ndim).astype(np.int32)
        m.dash_dir = np.tile(np.arange(ndim), side.size).astype(np.int32)
        m.dash_eta = np.full(side.size * ndim, 2.0e5)
    if load:
        top = bnd["top"]
        src = int(top[len(top) // 2])
        m.add_point_load(src, [0.4, -1.0] if ndim == 2 else [0.4, 0.2, -1.0],
                         ricker_velocity(np.arange(nt) * dt, 0.02, 30.0, 2e5))
    g = base_motion_series(nt, dt, length)
    for n in bnd["bottom"]:
        for d in dirs:
            m.add_support_motion(int(n), int(d), g, factor)
    return m, g, bnd
