"""Synthetic workloads of the dynamic hot path, built in SVL's conventions.

* ``layer2d``  -- 2D plane-strain elastic layer, lin2DQuad4 nx x ny elements of h m,
  rho = 2000, Vs = 200, nu = 0.25, Lysmer dashpots on left/right/bottom, vertical
  Ricker point load (velocity form, f0 = 5 Hz, t0 = 0.25 s, Ap = 1e5 N) at the
  surface centre node, CentralDifference with dt = 0.5 h / Vp. At 200 x 100 this is
  the 40,602-DOF, 2000-step layer.
* ``halfspace3d`` -- 3D elastic half-space of lin3DHexa8 elements of h m with six
  layers Vs 150 .. 800 m/s increasing with depth and a per-element lognormal
  perturbation exp(sigma z - sigma^2/2), z ~ N(0,1), sigma = 0.1, drawn with
  ``numpy.random.default_rng(seed)`` in element-tag order; nu = 0.3, rho = 2000;
  Lysmer dashpots on the bottom and four sides with per-node coefficients from the
  adjacent element's rho * V * h^2; vertical Ricker f0 = 3 Hz, t0 = 0.4 s at the
  surface centre; dt = 0.5 min_e(h / Vp_e). At 100 x 100 x 50 elements of 2 m this is
  the 1,560,753-DOF half-space.
* ``block3d_implicit`` -- 3D elastic block of lin3DHexa8 elements of h m, base fixed,
  per-element Vs ~ U[200, 1000] m/s drawn with ``numpy.random.default_rng(seed)`` in
  element-tag order, nu = 0.3, rho = 2000, Rayleigh damping of 5 % at 1 and 10 Hz
  (am = 2 zeta w1 w2 / (w1 + w2) = 0.5712, ak = 2 zeta / (w1 + w2) = 1.447e-3), lumped
  mass, vertical Ricker f0 = 5 Hz (t0 = 0.25 s, Ap = 1e6 N) at the top centre node,
  ExtendedNewmarkBeta with dt = 2e-3 s. At 64 x 64 x 64 this is the 811,200-DOF block
  of 500 implicit steps.
* ``halfspace3d_drm`` -- the half-space driven through a domain reduction layer instead of
  the point source (synthetic SV plane-wave signals).

Lysmer convention (Builder.py:1086-1131, per-node generalisation): a boundary node's
dashpot on the boundary-normal direction takes rho Vp A, the tangential ones rho Vs A,
with A = h * th in 2D (the convention of SVL's 2D plane-wave examples) and h^2 in 3D;
corners are counted once per boundary they lie on, as SVL does.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh import (ELEM_HEXA8, ELEM_QUAD4, Model, lysmer_dashpots, make_domain_area, make_domain_volume,
                   normal_dir, ricker_velocity, tags_for)


def layer2d(nx: int = 200, ny: int = 100, h: float = 1.0, nt: int = 2001, rho: float = 2000.0,
            vs: float = 200.0, nu: float = 0.25, f0: float = 5.0, t0: float = 0.25, Ap: float = 1.0e5,
            cfl: float = 0.5, lysmer: bool = True):
    """Returns (model, dt, nt)."""
    Lx, Ly = nx * h, ny * h
    P0 = (-Lx / 2.0, -Ly)
    coords, conn, bnd = make_domain_area(nx, ny, P0, (Lx / 2.0, -Ly), (-Lx / 2.0, 0.0))
    G = rho * vs * vs
    E = 2.0 * G * (1.0 + nu)
    vp = vs * math.sqrt(2.0 * (1.0 - nu) / (1.0 - 2.0 * nu))
    th = 1.0
    m = Model(ndim=2, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_QUAD4, conn=conn,
              E=E, nu=nu, rho=rho, thickness=th, name=f"layer2d_{nx}x{ny}")
    if lysmer:
        eta_s, eta_p = rho * vs * h * th, rho * vp * h * th
        names = ["left", "right", "bottom"]
        dn, dd, de = lysmer_dashpots(2, bnd, names,
                                     lambda n, name, k: eta_p if k == normal_dir(2, name) else eta_s)
        m.dash_node, m.dash_dir, m.dash_eta = dn, dd, de
    dt = cfl * h / vp
    t = np.arange(nt) * dt
    src = m.nearest_node((0.0, 0.0))
    m.add_point_load(src, [0.0, -1.0], ricker_velocity(t, t0, f0, Ap))
    return m, dt, nt


def halfspace3d(nx: int = 100, ny: int = 100, nz: int = 50, h: float = 2.0, nt: int = 4001,
                seed: int = 20260301, sigma: float = 0.1, nu: float = 0.3, rho: float = 2000.0,
                vs_layers=(150.0, 250.0, 350.0, 450.0, 600.0, 800.0), f0: float = 3.0, t0: float = 0.4,
                Ap: float = 1.0e6, cfl: float = 0.5, lysmer: bool = True):
    """Returns (model, dt, nt)."""
    Lx, Ly, Lz = nx * h, ny * h, nz * h
    P0 = (-Lx / 2.0, -Ly / 2.0, -Lz)
    coords, conn, bnd = make_domain_volume(nx, ny, nz, P0, (Lx / 2.0, -Ly / 2.0, -Lz),
                                           (-Lx / 2.0, Ly / 2.0, -Lz), (-Lx / 2.0, -Ly / 2.0, 0.0))
    ne = conn.shape[0]
    # Element depth below the surface -> layer index (equal-thickness layers).
    zc = coords[conn].mean(axis=1)[:, 2]
    depth = -zc
    nl = len(vs_layers)
    layer = np.minimum((depth / Lz * nl).astype(np.int64), nl - 1)
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(ne)
    vs = np.asarray(vs_layers)[layer] * np.exp(sigma * z - 0.5 * sigma * sigma)
    G = rho * vs * vs
    E = 2.0 * G * (1.0 + nu)
    vp = vs * math.sqrt(2.0 * (1.0 - nu) / (1.0 - 2.0 * nu))
    m = Model(ndim=3, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_HEXA8, conn=conn,
              E=E, nu=np.full(ne, nu), rho=np.full(ne, rho), name=f"halfspace3d_{nx}x{ny}x{nz}")
    if lysmer:
        # Adjacent element of each node: the first element (ascending tag) touching it.
        first = np.full(coords.shape[0], ne, dtype=np.int64)
        flat = conn.reshape(-1)
        np.minimum.at(first, flat, np.arange(flat.size) // conn.shape[1])
        area = h * h
        names = ["left", "right", "front", "back", "bottom"]

        def eta(n, name, k):
            e = first[n]
            v = vp[e] if k == normal_dir(3, name) else vs[e]
            return rho * v * area

        dn, dd, de = lysmer_dashpots(3, bnd, names, eta)
        m.dash_node, m.dash_dir, m.dash_eta = dn, dd, de
    dt = cfl * float(np.min(h / vp))
    t = np.arange(nt) * dt
    src = m.nearest_node((0.0, 0.0, 0.0))
    m.add_point_load(src, [0.0, 0.0, -1.0], ricker_velocity(t, t0, f0, Ap))
    return m, dt, nt


def halfspace3d_drm(nx: int = 100, ny: int = 100, nz: int = 50, h: float = 2.0, nt: int = 331,
                    box_frac: float = 0.8, f0: float = 3.0, t0: float = 0.4, amp: float = 1e-3, **kw):
    """The half-space of ``halfspace3d`` without its point source, loaded instead through a
    domain reduction layer: the box |x|, |y| <= box_frac L/2, z >= -box_frac Lz (open at the
    free surface) selected as SVL's pre-processor does, with synthetic node signals of a
    vertically incident SV Ricker pulse (x displacement amp r(t - t0 - (z + Lz)/vs) plus its
    surface reflection, vs = 150 m/s; V and A by finite differences in time). The signals
    are timing input, not the layered medium's exact free field. Returns (model, dt, nt)."""
    from .drm import attach, drm_box
    m, dt, nt = halfspace3d(nx, ny, nz, h=h, nt=nt, **kw)
    m.load_node = np.zeros(0, np.int32)
    m.load_dir, m.load_factor, m.load_series = np.zeros((0, 3)), np.zeros(0), np.zeros((0, 1))
    Lx, Ly, Lz = nx * h, ny * h, nz * h
    elems, nodes, cond = drm_box(m, (0.0, 0.0, 0.0), (box_frac * Lx / 2, box_frac * Ly / 2, box_frac * Lz))
    t = np.arange(nt) * dt
    z = m.coords[nodes, 2]
    vs = 150.0

    def ricker(tau):
        a = (math.pi * f0 * tau) ** 2
        return (1.0 - 2.0 * a) * np.exp(-a)

    u = amp * (ricker(t[None, :] - t0 - (z[:, None] + Lz) / vs) + ricker(t[None, :] - t0 - (Lz - z[:, None]) / vs))
    v = np.gradient(u, dt, axis=1)
    a = np.gradient(v, dt, axis=1)
    sig = np.zeros((nodes.size, nt, 9))
    sig[:, :, 0], sig[:, :, 3], sig[:, :, 6] = u, v, a
    attach(m, elems, nodes, cond, sig)
    m.name = f"halfspace3d_drm_{nx}x{ny}x{nz}"
    return m, dt, nt


def rayleigh(zeta: float, f1: float, f2: float):
    """(am, ak) giving damping ratio zeta at the frequencies f1 and f2 (Hz)."""
    w1, w2 = 2.0 * math.pi * f1, 2.0 * math.pi * f2
    return 2.0 * zeta * w1 * w2 / (w1 + w2), 2.0 * zeta / (w1 + w2)


def block3d_implicit(nx: int = 64, ny: int = 64, nz: int = 64, h: float = 1.0, nt: int = 501,
                     seed: int = 20260302, vs_range=(200.0, 1000.0), nu: float = 0.3, rho: float = 2000.0,
                     zeta: float = 0.05, f0: float = 5.0, t0: float = 0.25, Ap: float = 1.0e6, dt: float = 2e-3):
    """Returns (model, dt, nt)."""
    Lx, Ly, Lz = nx * h, ny * h, nz * h
    P0 = (-Lx / 2.0, -Ly / 2.0, -Lz)
    coords, conn, bnd = make_domain_volume(nx, ny, nz, P0, (Lx / 2.0, -Ly / 2.0, -Lz),
                                           (-Lx / 2.0, Ly / 2.0, -Lz), (-Lx / 2.0, -Ly / 2.0, 0.0))
    ne = conn.shape[0]
    rng = np.random.default_rng(seed)
    vs = rng.uniform(vs_range[0], vs_range[1], ne)
    E = 2.0 * rho * vs * vs * (1.0 + nu)
    restrained = np.zeros(coords.shape, bool)
    restrained[bnd["bottom"]] = True
    am, ak = rayleigh(zeta, 1.0, 10.0)
    m = Model(ndim=3, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_HEXA8, conn=conn,
              E=E, nu=np.full(ne, nu), rho=np.full(ne, rho), restrained=restrained, ray_am=np.full(ne, am),
              ray_ak=np.full(ne, ak), name=f"block3d_{nx}x{ny}x{nz}")
    t = np.arange(nt) * dt
    src = m.nearest_node((0.0, 0.0, 0.0))
    m.add_point_load(src, [0.0, 0.0, -1.0], ricker_velocity(t, t0, f0, Ap))
    return m, dt, nt


def soil_structure_j2(nx: int = 80, ny: int = 80, nz: int = 40, h: float = 2.0, nt: int = 1001,
                      block=(10, 10, 8), seed: int = 20260305, vs_layers=(150.0, 250.0, 400.0, 600.0),
                      nu: float = 0.3, rho: float = 2000.0, zeta: float = 0.05, f0: float = 3.0, t0: float = 0.4,
                      Ap: float = 2.0e8, dt: float = 2e-3):
    """Nonlinear soil-structure block: Hexa8 soil in len(vs_layers) equal layers of Plastic3DJ2
    (top layer first; K, G from Vs and nu; sigma_y = sqrt(3) c with the cohesion c ~ U[50, 150]
    kPa per layer from default_rng(seed); H = 0.1 G; beta = 1) around an embedded elastic block
    of block = (bx, by, bz) elements at the surface centre (E = 30 GPa, nu = 0.2, rho = 2400).
    Base fixed, Rayleigh damping zeta at 1 and 10 Hz, a horizontal Ricker point load (velocity
    form) on the block's top centre. Returns (model, dt, nt)."""
    Lx, Ly, Lz = nx * h, ny * h, nz * h
    P0 = (-Lx / 2.0, -Ly / 2.0, -Lz)
    coords, conn, bnd = make_domain_volume(nx, ny, nz, P0, (Lx / 2.0, -Ly / 2.0, -Lz),
                                           (-Lx / 2.0, Ly / 2.0, -Lz), (-Lx / 2.0, -Ly / 2.0, 0.0))
    ne = conn.shape[0]
    kk, rem = np.divmod(np.arange(ne), nx * ny)
    jj, ii = np.divmod(rem, nx)
    nl = len(vs_layers)
    layer = np.minimum((nz - 1 - kk) * nl // nz, nl - 1)  # 0 = top layer
    rng = np.random.default_rng(seed)
    coh = rng.uniform(50e3, 150e3, nl)
    vs = np.asarray(vs_layers, dtype=np.float64)[layer]
    G = rho * vs * vs
    K = 2.0 * G * (1.0 + nu) / (3.0 * (1.0 - 2.0 * nu))
    E = 2.0 * G * (1.0 + nu)
    bx, by, bz = block
    inb = ((ii >= (nx - bx) // 2) & (ii < (nx - bx) // 2 + bx) & (jj >= (ny - by) // 2) & (jj < (ny - by) // 2 + by)
           & (kk >= nz - bz))
    E = np.where(inb, 30.0e9, E)
    nue = np.where(inb, 0.2, nu)
    rhoe = np.where(inb, 2400.0, rho)
    restrained = np.zeros(coords.shape, bool)
    restrained[bnd["bottom"]] = True
    am, ak = rayleigh(zeta, 1.0, 10.0)
    m = Model(ndim=3, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_HEXA8, conn=conn,
              E=E, nu=nue, rho=rhoe, restrained=restrained, ray_am=np.full(ne, am), ray_ak=np.full(ne, ak),
              name=f"soil_structure_{nx}x{ny}x{nz}")
    soil = np.nonzero(~inb)[0]
    m.set_j2(K[soil], G[soil], np.sqrt(3.0) * coh[layer[soil]], 0.1 * G[soil], 1.0, soil)
    t = np.arange(nt) * dt
    src = m.nearest_node((0.0, 0.0, 0.0))
    m.add_point_load(src, [1.0, 0.0, 0.0], ricker_velocity(t, t0, f0, Ap))
    return m, dt, nt


def realization2d(nx: int, ny: int, seed: int, h: float = 1.0, nt: int = 3001, vs_mean: float = 300.0,
                  sigma: float = 0.2, nu: float = 0.25, rho: float = 2000.0, f0: float | None = None,
                  cfl: float = 0.5, x0: float = 0.0):
    """One 2D plane-strain realization of the batched workload: lin2DQuad4 nx x ny elements of
    h m with its left edge at x0, per-element Vs lognormal with mean vs_mean,
    Vs = vs_mean exp(sigma z - sigma^2/2), z ~ N(0,1) from ``default_rng(seed)`` in element-tag
    order (drawn after f0 ~ U[2, 8] Hz when f0 is None); Lysmer dashpots on left/right/bottom
    (per node from the first adjacent element, rho V h th); a vertically incident SV plane
    wave as Lysmer-consistent base forces 2 eta_s(node) v_in(t) in x on the bottom nodes, with
    v_in the Ricker velocity waveform (t0 = 1.2/f0, unit amplitude); CentralDifference with
    dt = cfl min_e(h / Vp_e). Returns (model, dt, nt)."""
    rng = np.random.default_rng(seed)
    if f0 is None:
        f0 = float(rng.uniform(2.0, 8.0))
    Lx, Ly = nx * h, ny * h
    coords, conn, bnd = make_domain_area(nx, ny, (x0, -Ly), (x0 + Lx, -Ly), (x0, 0.0))
    ne = conn.shape[0]
    z = rng.standard_normal(ne)
    vs = vs_mean * np.exp(sigma * z - 0.5 * sigma * sigma)
    G = rho * vs * vs
    E = 2.0 * G * (1.0 + nu)
    vp = vs * math.sqrt(2.0 * (1.0 - nu) / (1.0 - 2.0 * nu))
    th = 1.0
    m = Model(ndim=2, coords=coords, node_tags=tags_for(coords.shape[0]), elem_type=ELEM_QUAD4, conn=conn,
              E=E, nu=np.full(ne, nu), rho=np.full(ne, rho), thickness=th, name=f"realization2d_{nx}x{ny}_{seed}")
    first = np.full(coords.shape[0], ne, dtype=np.int64)
    flat = conn.reshape(-1)
    np.minimum.at(first, flat, np.arange(flat.size) // conn.shape[1])

    def eta(n, name, k):
        e = first[n]
        return rho * (vp[e] if k == normal_dir(2, name) else vs[e]) * h * th

    dn, dd, de = lysmer_dashpots(2, bnd, ["left", "right", "bottom"], eta)
    m.dash_node, m.dash_dir, m.dash_eta = dn, dd, de
    dt = cfl * float(np.min(h / vp))
    t = np.arange(nt) * dt
    vin = ricker_velocity(t, 1.2 / f0, f0, 1.0)
    base = np.sort(bnd["bottom"]).astype(np.int32)
    m.load_node = base
    m.load_dir = np.ascontiguousarray(np.tile([1.0, 0.0], (base.size, 1)))
    m.load_factor = np.array([2.0 * eta(int(n), "bottom", 0) for n in base])
    m.load_series = np.ascontiguousarray(np.tile(vin, (base.size, 1)))
    return m, dt, nt


def batch2d(nreal: int = 64, seed0: int = 20260304, nt: int = 3001, nx_range=(50, 400), ny_range=(50, 200),
            gap: int = 8):
    """The batched workload: realization i uses ``default_rng(seed0 + i)`` for nx ~ U{nx_range},
    ny ~ U{ny_range}, f0 ~ U[2, 8] Hz and its Vs field (realization2d); realizations are laid
    side by side along x with ``gap`` element widths between them. Returns (models, dts, nt)."""
    models, dts = [], []
    x0 = 0.0
    for i in range(nreal):
        rng = np.random.default_rng(seed0 + i)
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        f0 = float(rng.uniform(2.0, 8.0))
        m, dt, _ = realization2d(nx, ny, seed0 + i, nt=nt, f0=f0, x0=x0)
        models.append(m)
        dts.append(dt)
        x0 += nx + gap
    return models, dts, nt


def dofs_of(model: Model, node: int):
    """Free DOF indices of a node (restrained components are skipped)."""
    f = model.free_dof()[node]
    return [int(x) for x in f if x >= 0]
