"""Equilibrium-tide body force on the sphere (swe_set_tide; no reference
counterpart, so pinned by its closed form and by a balance):

* the term the residual gains is exactly g h grad(eta_eq) of the stated
  potential at the stage time (probe residual with and without it);
* a resting ocean whose surface IS a static equilibrium tide stays at rest to
  the scheme's truncation error, while the opposite sign would double the
  imbalance: this fixes sign and magnitude;
* batched basins with tides are bit-identical to their plain runs;
* the adjoint differentiates through it (central finite differences)."""
import math

import numpy as np
import pytest

import swe_amd

pytestmark = pytest.mark.gpu

G = 9.80665
R_SMALL = 2.0e6            # a small planet keeps a coarse grid's cells of modest aspect


def sphere(nx, ny, lat0, lat1, *, R=R_SMALL, manning=0.0, order=2, limiter="mc", omega=7.2921e-5):
    rows = swe_amd.geometry_spherical(nx, ny, R=R, lat0=lat0, lat1=lat1, omega=omega)
    s = swe_amd.Solver(nx, ny, rows, sphere=True, R=R, manning=manning, order=order, limiter=limiter)
    lat, lon = swe_amd.sphere_lonlat(nx, ny, lat0, lat1)
    return s, rows, lat, lon


def eta_eq(cons, lat, lon, t):
    LA, LO = np.meshgrid(lat, lon, indexing="ij")
    e = np.zeros_like(LA)
    for m, A, om, ph in cons:
        f = np.cos(LA) ** 2 if m == 2 else np.sin(2 * LA)
        e += A * f * np.cos(om * t + ph + m * LO)
    return e


def test_tide_term_is_g_h_grad_eta_eq():
    nx, ny = 48, 24
    s, rows, lat, lon = sphere(nx, ny, -60.0, 60.0)
    W, H = nx + 4, ny + 4
    rng = np.random.default_rng(1)
    b = -1000.0 + rng.normal(0, 20, W * H)
    h = -b + rng.normal(0, 0.5, W * H)
    hu = h * rng.normal(0, 0.1, W * H)
    hv = h * rng.normal(0, 0.1, W * H)
    s.load_state(b, h, hu, hv)
    s.set_time(3600.0)
    cons = [(2, 0.5, swe_amd.TIDE_M2[1], 0.3), (2, 0.2, swe_amd.TIDE_S2[1], 1.1), (1, 0.3, swe_amd.TIDE_K1[1], -0.4)]
    dt = 20.0
    s.residual(dt)
    base = [s.get_field(f).copy() for f in (swe_amd.F_RH, swe_amd.F_RHU, swe_amd.F_RHV)]
    s.set_tide(cons, lat, lon)
    s.residual(dt)
    tide = [s.get_field(f).copy() for f in (swe_amd.F_RH, swe_amd.F_RHU, swe_amd.F_RHV)]
    assert np.array_equal(tide[0], base[0])            # mass is untouched
    LA, LO = np.meshgrid(lat, lon, indexing="ij")
    fx = np.zeros_like(LA)
    fy = np.zeros_like(LA)
    t = 3600.0
    for m, A, om, ph in cons:
        th = om * t + ph + m * LO
        if m == 2:     # eta = A cos^2 cos(th): (1/(R cos)) d/dlon, (1/R) d/dlat
            fx += -2 * A * np.cos(LA) * np.sin(th) / R_SMALL
            fy += -A * np.sin(2 * LA) * np.cos(th) / R_SMALL
        else:          # eta = A sin(2 lat) cos(th)
            fx += -2 * A * np.sin(LA) * np.sin(th) / R_SMALL
            fy += 2 * A * np.cos(2 * LA) * np.cos(th) / R_SMALL
    inner = np.zeros((H, W), bool)
    inner[2:-2, 2:-2] = True
    gx = (G * h.reshape(H, W) * fx)[inner]
    gy = (G * h.reshape(H, W) * fy)[inner]
    dx_ = (tide[1] - base[1]).reshape(H, W)[inner]
    dy_ = (tide[2] - base[2]).reshape(H, W)[inner]
    scale = max(np.abs(gx).max(), np.abs(gy).max())
    assert np.abs(dx_ - gx).max() <= 1e-12 * scale
    assert np.abs(dy_ - gy).max() <= 1e-12 * scale
    # removing it restores the residual bit for bit
    s.set_tide([], lat, lon)
    s.residual(dt)
    assert np.array_equal(s.get_field(swe_amd.F_RHU), base[1])


@pytest.mark.parametrize("m", [2, 1])
def test_static_equilibrium_tide_is_balanced(m):
    # central slopes: no limiter flattens the reconstruction at the extrema of the
    # surface, so what is left is the scheme's O(dx^2) truncation error
    nx, ny = 96, 48
    s, rows, lat, lon = sphere(nx, ny, -60.0, 60.0, omega=0.0, limiter="central")
    W, H = nx + 4, ny + 4
    cons = [(m, 0.5, 0.0, 0.0)]                       # a frozen potential
    e = eta_eq(cons, lat, lon, 0.0).ravel()
    b = np.full(W * H, -1000.0)
    h = e - b
    z = np.zeros(W * H)
    s.load_state(b, h, z, z)
    s.apply_bc()
    s.set_tide(cons, lat, lon)
    s.residual(10.0)
    # away from the walls at +-60 degrees, where the mirrored ghost rows cannot
    # carry the potential's gradient into the wall cells
    inner = np.zeros((H, W), bool)
    inner[5:-5, 2:-2] = True
    inner = inner.ravel()
    rhu, rhv = s.get_field(swe_amd.F_RHU)[inner], s.get_field(swe_amd.F_RHV)[inner]
    # the forcing alone, for scale: g h |grad eta_eq|
    s.set_tide([], lat, lon)
    s.residual(10.0)
    pu, pv = s.get_field(swe_amd.F_RHU)[inner], s.get_field(swe_amd.F_RHV)[inner]
    force = max(np.abs(pu).max(), np.abs(pv).max())
    imbalance = max(np.abs(rhu).max(), np.abs(rhv).max())
    print(f"  m={m}: pressure gradient {force:.4e}, residual with the tide {imbalance:.4e}")
    assert imbalance < 0.02 * force


def test_batched_tides_match_plain_runs():
    specs = [(36, 20, -60.0, 60.0, 0.02), (48, 30, -70.0, 50.0, 0.03), (24, 12, -40.0, 40.0, 0.015)]
    cons = [(2, 0.5, swe_amd.TIDE_M2[1], 0.0), (2, 0.2, swe_amd.TIDE_S2[1], 0.5), (1, 0.3, swe_amd.TIDE_K1[1], 1.0)]
    basins, states, plains = [], [], []
    for q, (nx, ny, l0, l1, n) in enumerate(specs):
        rows = swe_amd.geometry_spherical(nx, ny, R=R_SMALL, lat0=l0, lat1=l1, omega=7.2921e-5)
        W, H = nx + 4, ny + 4
        rng = np.random.default_rng(10 + q)
        b = -2000.0 + rng.normal(0, 50, W * H)
        h = np.maximum(0.0, -b)
        z = np.zeros(W * H)
        states.append((b, h, z, z.copy()))
        basins.append(dict(nx=nx, ny=ny, rows=rows, sphere=True, R=R_SMALL, manning=n))
        lat, lon = swe_amd.sphere_lonlat(nx, ny, l0, l1)
        s = swe_amd.Solver(nx, ny, rows, sphere=True, R=R_SMALL, manning=n)
        s.load_state(*states[-1])
        s.set_tide(cons, lat, lon)
        s.advance(30)
        plains.append((s.h.copy(), s.hu.copy(), s.hv.copy(), s.t))
    B = swe_amd.Batch(basins)
    for q, (nx, ny, l0, l1, n) in enumerate(specs):
        B.load_state(q, *states[q])
        lat, lon = swe_amd.sphere_lonlat(nx, ny, l0, l1)
        B.set_tide(q, cons, lat, lon)
    B.advance(30)
    for q in range(len(specs)):
        _, h, hu, hv = B.state(q)
        ph, phu, phv, pt = plains[q]
        assert B.get_scalar(q, swe_amd.S_T) == pt
        assert np.array_equal(h, ph) and np.array_equal(hu, phu) and np.array_equal(hv, phv), q
        assert np.abs(hu).max() > 0                     # the tide moved the water


@pytest.mark.parametrize("limiter,eps,tol", [("central", 1e-3, 1e-6), ("mc", 1e-6, 1e-4)])
def test_adjoint_through_the_tide(limiter, eps, tol):
    # Central slopes have no kink for a perturbation to cross, so the step can be
    # large (1e-5 m of surface), well above the forward's rounding noise (b + h
    # with |b|, |h| ~ 300 m); with MC a step that large lets limiters change
    # branch, so it stays small and the tolerance covers the noise.
    nx, ny = 36, 20
    s, rows, lat, lon = sphere(nx, ny, -50.0, 50.0, limiter=limiter)
    W, H = nx + 4, ny + 4
    rng = np.random.default_rng(3)
    b = -300.0 + rng.normal(0, 5, W * H)
    h0 = -b + rng.normal(0, 0.05, W * H)
    z = np.zeros(W * H)
    s.load_state(b, h0, z, z)
    s.set_tide([(2, 2.0, swe_amd.TIDE_M2[1], 0.2), (1, 1.0, swe_amd.TIDE_K1[1], 0.7)], lat, lon)
    nf = rng.uniform(0.015, 0.035, W * H)
    nsteps, every = 40, 4
    gauges = np.array([(4 + 3 * q) * W + 5 + 7 * q for q in range(5)], dtype=np.int32)
    adj = swe_amd.Adjoint(s, nsteps)
    adj.set_gauges(gauges, every, np.zeros((nsteps // every, len(gauges))))
    adj.forward(h0, nf)
    obs = 0.5 * adj.series() + 0.01
    adj.set_gauges(gauges, every, obs)
    adj.forward(h0, nf)
    gh, gn = adj.backward()
    wet = np.zeros((H, W), bool)
    wet[2:-2, 2:-2] = True
    wet = wet.ravel()
    X, Y = np.meshgrid(np.arange(W), np.arange(H))
    dh = (np.sin(X / 4.0) * np.cos(Y / 3.0)).ravel() * wet * 0.01
    fd = (adj.loss(h0 + eps * dh, nf) - adj.loss(h0 - eps * dh, nf)) / (2 * eps)
    ad = float(np.dot(gh, dh))
    assert abs(ad - fd) <= tol * abs(fd), (ad, fd)
    dn = rng.uniform(-1, 1, W * H) * wet * 1e-3
    fdn = (adj.loss(h0, nf + 1e-3 * dn) - adj.loss(h0, nf - 1e-3 * dn)) / 2e-3
    adn = float(np.dot(gn, dn))
    assert abs(adn - fdn) <= max(tol, 1e-6) * abs(fdn), (adn, fdn)
