"""Pins the oracle's J2 plasticity and Newton-Raphson restatement (CPU).

Closed forms on one Gauss point (oracle.j2_path, one UpdateState per strain):
* pure shear with linear hardening: tau = G gamma up to tau_y = sigma_y / sqrt(3), then slope
  G H / (H + 3 G), for isotropic (beta = 1) and kinematic (beta = 0) hardening, with any
  step size (radial loading);
* reversal: isotropic hardening re-yields at -tau_max, kinematic at tau_max - 2 tau_y;
* the consistent tangent equals the finite-difference derivative of the stress of one update
  (3D and plane strain);
* SVL's iterate accumulation: a second UpdateState at a smaller strain starts from the first
  call's plastic state (elastic unloading from it), not from the committed state.
SVL's validation cases F03 (one J2 Quad4, Newton) and F07 (J2 soil column, Newton) against
the OpenSees histories they ship: the oracle's response follows them to the accuracy the
cases themselves reach (OpenSees' J2Plasticity is a different return mapping), asserted at
a bound a dropped hardening term or a wrong plane-strain index would exceed.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from svl_amd.mesh import ELEM_QUAD4, Model, tags_for

GOLD = Path(__file__).resolve().parent / "golden"
K, G, SY, H = 1.333333e8, 8.0e7, 4.0e7, 8.0e7


def shear_path(gammas):
    e = np.zeros((len(gammas), 6))
    e[:, 3] = gammas
    return e


def tau_monotonic(g, beta):
    ty = SY / np.sqrt(3.0)
    gy = ty / G
    return np.where(g <= gy, G * g, ty + G * H / (H + 3 * G) * (g - gy))


@pytest.mark.parametrize("beta", [1.0, 0.0])
def test_pure_shear_monotonic(beta):
    gy = SY / np.sqrt(3.0) / G
    for gam in (np.linspace(0, 6 * gy, 31)[1:], np.array([0.5 * gy, 6 * gy])):
        sig, tan, al = oracle.j2_path([K, G, SY, H, beta], shear_path(gam))
        assert np.allclose(sig[:, 3], tau_monotonic(gam, beta), rtol=1e-12, atol=1e-6)
        assert np.allclose(sig[:, [0, 1, 2, 4, 5]], 0.0, atol=1e-6)
        # Plastic tangent in shear = G H / (H + 3G).
        assert abs(tan[-1, 3, 3] - G * H / (H + 3 * G)) < 1e-9 * G


@pytest.mark.parametrize("beta", [1.0, 0.0])
def test_shear_reversal(beta):
    ty = SY / np.sqrt(3.0)
    gy = ty / G
    ga = 4 * gy
    down = np.linspace(ga, -6 * gy, 201)
    sig, _, _ = oracle.j2_path([K, G, SY, H, beta], shear_path(np.concatenate([[ga], down[1:]])))
    tau = sig[:, 3]
    ta = tau[0]
    reyield = -ta if beta == 1.0 else ta - 2 * ty
    el = G * (down[1:] - ga) + ta
    elastic = el >= reyield
    assert np.allclose(tau[1:][elastic], el[elastic], rtol=1e-12, atol=1e-3)
    # Past re-yield: tau = reyield + G_t (gamma - gamma_ry), G_t = G H / (H + 3 G).
    g_ry = ga + (reyield - ta) / G
    plastic = down[1:][~elastic]
    assert plastic.size > 50
    assert np.allclose(tau[1:][~elastic], reyield + G * H / (H + 3 * G) * (plastic - g_ry), rtol=1e-10)


@pytest.mark.parametrize("plane", [False, True])
def test_tangent_is_stress_derivative(plane):
    rng = np.random.default_rng(3)
    nv = 3 if plane else 6
    gy = SY / np.sqrt(3.0) / G
    path = np.cumsum(rng.standard_normal((6, nv)) * 2 * gy, axis=0)
    path = np.vstack([path, 1.5 * path[-1]])  # a last step outward from the current strain
    p = [K, G, SY, H, 0.4]
    _, tan, al = oracle.j2_path(p, path, plane=plane)
    assert al[-1] > al[-2]  # the last update is plastic
    h = 1e-9
    fd = np.zeros((nv, nv))
    for j in range(nv):
        pp, pm = path.copy(), path.copy()
        pp[-1, j] += h
        pm[-1, j] -= h
        fd[:, j] = (oracle.j2_path(p, pp, plane=plane)[0][-1] - oracle.j2_path(p, pm, plane=plane)[0][-1]) / (2 * h)
    assert np.allclose(tan[-1], fd, rtol=1e-5, atol=1e-5 * G)


def test_iterate_accumulation():
    ty = SY / np.sqrt(3.0)
    gy = ty / G
    ga, gb = 4 * gy, 3 * gy
    sig, _, _ = oracle.j2_path([K, G, SY, H, 1.0], shear_path([ga, gb]))
    ta = tau_monotonic(np.array([ga]), 1.0)[0]
    # SVL: the second call starts from the first call's plastic strain -> elastic unloading.
    assert abs(sig[1, 3] - (ta - G * (ga - gb))) < 1e-9 * ta
    # The committed-state answer would be on the monotonic curve instead.
    assert abs(sig[1, 3] - tau_monotonic(np.array([gb]), 1.0)[0]) > 1e-3 * ta


def load(name):
    with np.load(GOLD / f"{name}.npz") as z:
        return {k: z[k] for k in z.files}


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def test_F03_j2_quad4_newton():
    g = load("F03")
    coords = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0]])
    restrained = np.array([[True, True], [False, True], [False, True], [False, True]])
    m = Model(ndim=2, coords=coords, node_tags=tags_for(4), elem_type=ELEM_QUAD4, conn=np.array([[0, 1, 2, 3]]),
              E=1.0, nu=0.3, rho=0.0, thickness=1.0, restrained=restrained)
    m.set_j2(1.333333e8, 8.0e7, 4.0e7, 8.0e7, 1.0)
    m.add_point_load(2, [8.0e7, 0.0], g["load"])
    rec = [int(m.free_dof()[2][0])]
    r = oracle.run(m, oracle.NEWMARK, 0.001, 1000, rec, algorithm=oracle.NEWTON, nmax=100, cnvgtest=5, cnvgtol=1e-6)
    os_u = g["displacement"][:, 1]
    err = rel_rms(r["U"][:, 0], os_u)
    assert err < 2e-3, err
    assert r["iters"].max() > 1  # the material yields and Newton iterates
    # Kinematic instead of isotropic hardening (beta = 0) departs visibly (measured 0.73).
    m.set_j2(1.333333e8, 8.0e7, 4.0e7, 8.0e7, 0.0)
    r0 = oracle.run(m, oracle.NEWMARK, 0.001, 1000, rec, algorithm=oracle.NEWTON, nmax=100, cnvgtest=5,
                    cnvgtol=1e-6)
    assert rel_rms(r0["U"][:, 0], os_u) > 0.1


def test_F07_j2_soil_column_newton():
    g = load("F07")
    n = 101
    coords = np.zeros((2 * n, 2))
    coords[0::2, 0], coords[1::2, 0] = 0.0, 1.0
    coords[0::2, 1] = coords[1::2, 1] = np.arange(n, dtype=float)
    k = np.arange(n - 1)
    conn = np.stack([2 * k, 2 * k + 1, 2 * k + 3, 2 * k + 2], axis=1).astype(np.int32)
    restrained = np.zeros(coords.shape, bool)
    restrained[:, 1] = True
    m = Model(ndim=2, coords=coords, node_tags=tags_for(2 * n), elem_type=ELEM_QUAD4, conn=conn, E=1.0, nu=0.3,
              rho=2000.0, thickness=1.0, restrained=restrained, ray_am=0.0, ray_ak=0.0005,
              dash_node=np.array([0, 1]), dash_dir=np.array([0, 0]), dash_eta=np.array([1.0e6, 1.0e6]))
    nt = 2001
    m.add_point_load(0, [1.0e6, 0.0], g["load"][:nt])
    m.add_point_load(1, [1.0e6, 0.0], g["load"][:nt])
    rec = [int(m.free_dof()[0][0]), int(m.free_dof()[200][0])]
    os_u = g["displacement"][:nt - 1, [1, 3]]

    def run(beta):
        m.set_j2(2.9e7, 2.0e7, 1.0e4, 1.0e7, beta)
        return oracle.run(m, oracle.NEWMARK, 0.001, nt, rec, algorithm=oracle.NEWTON, nmax=50, cnvgtest=5,
                          cnvgtol=1e-6)

    r = run(1.0)
    err = rel_rms(r["U"], os_u)
    assert err < 1e-4, err      # measured 1.7e-6
    assert r["iters"].max() > 2  # yielding: Newton takes several iterations
    assert rel_rms(run(0.0)["U"], os_u) > 1e-2  # kinematic hardening instead: 8e-2
