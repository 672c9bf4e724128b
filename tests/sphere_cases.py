"""The spherical cases of the reference's tools/verify-sphere.mjs whose results
the reference records in tools/sphere-baseline.json (copied, as data, to
tests/golden/sphere_baseline.json): the area identity (:229-249), the resting
ocean over three beds (:480-521) and the Merian free oscillation (:673-746),
restated for the GPU solver and the oracle."""
import json
import math
import os

import numpy as np

import swe_amd

HERE = os.path.dirname(os.path.abspath(__file__))
SPHERE_BASELINE_JSON = os.path.join(HERE, "golden", "sphere_baseline.json")
G_REF, R_REF, H_OCEAN = 9.80665, 6371e3, 4000.0
D = math.pi / 180


def baseline():
    with open(SPHERE_BASELINE_JSON) as f:
        return json.load(f)


def wrap180(x):
    return math.fmod(math.fmod(x + 180, 360) + 360, 360) - 180


def bed_flat(lon, lat):
    return -H_OCEAN


def bed_uneven(lon, lat):
    return -H_OCEAN + 900 * math.sin(3 * lon * D) * math.cos(2 * lat * D) + 350 * math.sin(5 * lat * D)


def bed_piercing(lon, lat):
    dl, dp = wrap180(lon - 30), lat - 20
    return bed_uneven(lon, lat) + 5000 * math.exp(-(dl * dl + dp * dp) / 128)


BEDS = {"flat": bed_flat, "uneven": bed_uneven, "piercing": bed_piercing}


def legendre(n, x):
    """The recurrence of verify-sphere.mjs:666-671."""
    if n == 0:
        return 1.0
    p0, p1 = 1.0, x
    for k in range(1, n):
        p0, p1 = p1, ((2 * k + 1) * x * p1 - k * p0) / (k + 1)
    return p1


def grid(deg, lat0=-90.0, lat1=90.0):
    nx, ny = round(360 / deg), round((lat1 - lat0) / deg)
    rows = swe_amd.geometry_spherical(nx, ny, R=R_REF, lat0=lat0, lat1=lat1, omega=0.0)
    lat, _ = swe_amd.sphere_lonlat(nx, ny, lat0, lat1)
    return nx, ny, rows, lat


def lonlat_deg(nx, lat, ng=2):
    """cellLonLat (swe.mjs:262-267) of every column and row, in degrees."""
    dlam = 2 * math.pi / nx
    lons = []
    for i in range(nx + 2 * ng):
        lon = (i - ng + 0.5) * dlam * 180 / math.pi
        lons.append(math.fmod(math.fmod(lon, 360) + 360, 360) - 180)
    lats = [p * 180 / math.pi for p in lat]
    return lons, lats


def sample(nx, ny, lat, bed, eta0):
    """The constructor's b and h (swe.mjs:222-231) for closures of (lon, lat) in degrees."""
    lons, lats = lonlat_deg(nx, lat)
    W, H = nx + 4, ny + 4
    b = np.empty(W * H)
    h = np.empty(W * H)
    for j in range(H):
        for i in range(W):
            bb = bed(lons[i], lats[j])
            b[j * W + i] = bb
            h[j * W + i] = max(0.0, eta0(lons[i], lats[j]) - bb)
    return b, h


def interior_area(rows, nx, ny, ng=2):
    """interiorArea (geometry.mjs:325-329): rows in order, each area x nx."""
    s = 0.0
    for j in range(ng, ny + ng):
        s += rows["area"][j] * nx
    return s


def resting_ocean(make, case, deg, order=2, lat0=-90.0, lat1=90.0, steps=400):
    """verify-sphere.mjs:497-521 on a solver built by make(nx, ny, rows, order): the
    dry cells at rest, then max speed, max |eta| and volume drift after `steps`
    steps of the initial CFL step."""
    nx, ny, rows, lat = grid(deg, lat0, lat1)
    b, h = sample(nx, ny, lat, BEDS[case], lambda lon, la: 0.0)
    sim = make(nx, ny, rows, order)
    z = np.zeros_like(h)
    sim.load_state(b, h, z, z)
    sim.set_bc(("periodic", "periodic", "reflect", "reflect"))
    W = nx + 4
    inner = np.zeros((ny + 4, W), bool)
    inner[2:-2, 2:-2] = True
    dry = int((h.reshape(ny + 4, W)[inner] <= 1e-3).sum())
    v0 = sim.volume()
    dt = sim.max_dt()
    for _ in range(steps):
        sim.step(dt)
    return dict(dry=dry, maxSpeed=sim.max_speed(), maxEta=sim.max_surface_deviation(0.0),
                drift=abs(sim.volume() - v0) / v0)


def merian_problem(n, deg):
    """verify-sphere.mjs:673-689: the flat 4000 m aquaplanet with eta = P_n(sin lat),
    the projection weights w_j = P_n(lat_j) area_j, and the closed-form period."""
    omega = math.sqrt(G_REF * H_OCEAN * n * (n + 1) / (R_REF * R_REF))
    t_closed = 2 * math.pi / omega
    nx, ny, rows, lat = grid(deg)
    b, h = sample(nx, ny, lat, bed_flat, lambda lon, la: legendre(n, math.sin(la * D)))
    _, lats = lonlat_deg(nx, lat)
    w = np.array([legendre(n, math.sin(lats[j + 2] * D)) * rows["area"][j + 2] for j in range(ny)])
    return dict(nx=nx, ny=ny, rows=rows, b=b, h=h, w=w, t_closed=t_closed)


def projection(w, b, h, nx, ny):
    """verify-sphere.mjs:686-689: sum_j (row sum of eta) w_j, in order (cumsum adds sequentially)."""
    W = nx + 4
    eta = (b + h).reshape(ny + 4, W)[2:-2, 2:-2]
    rows = np.cumsum(eta, axis=1)[:, -1]
    return float(np.cumsum(rows * w)[-1])


def merian_period(t_closed, times, vals, c0):
    """Zero crossings and the period of verify-sphere.mjs:690-700."""
    zeros = []
    prev, tprev = c0, 0.0
    for t, c in zip(times, vals):
        if (prev > 0 and c <= 0) or (prev < 0 and c >= 0):
            zeros.append(tprev + (t - tprev) * prev / (prev - c))
        prev, tprev = c, t
    T = 2 * (zeros[-1] - zeros[0]) / (len(zeros) - 1)
    return T, 100 * (T - t_closed) / t_closed, len(zeros)


def fixed_steps(dt, t_end):
    """Steps of `while (t < tEnd) step(dt)` with t accumulated as the solver does."""
    t, n = 0.0, 0
    while t < t_end:
        t += dt
        n += 1
    return n
