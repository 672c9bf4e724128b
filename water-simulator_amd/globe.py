"""The globe harness of the reference (src/globe.mjs) on the HIP solver: the
spherical mode-period measurement (measureModePeriod, globe.mjs:1387-1456) and
the polar timestep tax (polarTimestepTax, :1022-1037), for grids far finer than
the reference's host loop reaches, and for latitude-capped domains.

measure_mode_period runs the reference's loop -- step(min(maxDt, tEnd - t))
until tEnd, the area-weighted projection of eta onto P_n after every step, the
period from linearly interpolated zero crossings -- device-resident: the steps
run as one DT_UNTIL advance and the projection of every step is recorded on the
device (swe_set_projection) in the reference's summation order; the crossings
are found on the host from the recorded series.

    r = globe.measure_mode_period(n=2, deg=1, g_ref=9.80665)
    r["measured"], r["closed"], r["rel"]
"""
import math

import numpy as np

import swe_amd

D2R = math.pi / 180
R2D = 180 / math.pi
G_LOCAL = 9.80665


def js_round(x):
    """Math.round: the nearest integer, halves toward +infinity."""
    f = math.floor(x)
    return int(f + 1) if x - f >= 0.5 else int(f)


def legendre_p(n, x):
    """P_n(x) for n = 2, 3, 4 as globe.mjs:1088-1094 writes it out."""
    if n == 2:
        return 0.5 * (3 * x * x - 1)
    if n == 3:
        return 0.5 * (5 * x * x * x - 3 * x)
    if n == 4:
        return (35 * swe_amd.lib().swe_js_pow(x, 4.0) - 30 * x * x + 3) / 8
    raise ValueError(f"legendre_p: n = {n} is not one of 2, 3, 4 (n = 1 is a rigid translation of the shell, "
                     "n = 0 a change of volume)")


def polar_timestep_tax(rows, lat, cfl=0.45, H=4000.0, ng=2):
    """globe.mjs:1022-1037 off a grid's own metric rows: the CFL step a uniform
    depth-H ocean allows in the row closest to the equator and in the row with
    the smallest step, their ratio and latitudes (degrees).  lat: phiC per row."""
    c = math.sqrt(G_LOCAL * H)
    dx, dyc = rows["dxRow"], rows["dyCFL"]

    def dt_of(j):
        return cfl / (c / dx[j] + c / dyc[j])

    nrows = len(dx)
    eq_j, best = ng, math.inf
    for j in range(ng, nrows - ng):
        a = abs(lat[j])
        if a < best:
            best, eq_j = a, j
    worst_j = ng
    for j in range(ng, nrows - ng):
        if dt_of(j) < dt_of(worst_j):
            worst_j = j
    eq, pole = dt_of(eq_j), dt_of(worst_j)
    return dict(eq=eq, pole=pole, ratio=eq / pole, pole_lat=lat[worst_j] * R2D, eq_lat=lat[eq_j] * R2D)


def mode_problem(n=2, deg=6, H=4000.0, R=6371e3, amp=1.0, lat0=-90.0, lat1=90.0):
    """The grid, initial state and projection weights measureModePeriod builds
    (globe.mjs:1396-1415): flat bed -H, eta = amp P_n(sin lat), no rotation."""
    nx, ny = js_round(360 / deg), js_round((lat1 - lat0) / deg)
    rows = swe_amd.geometry_spherical(nx, ny, R=R, lat0=lat0, lat1=lat1, omega=0.0)
    lat, _ = swe_amd.sphere_lonlat(nx, ny, lat0, lat1)
    W, Hh, ng = nx + 4, ny + 4, 2
    # the constructor samples eta0(lon, lat) with lat in degrees (swe.mjs:265-266)
    eta_row = np.array([amp * legendre_p(n, math.sin((p * 180 / math.pi) * D2R)) for p in lat])
    b = np.full(W * Hh, -H)
    h = np.maximum(0.0, np.repeat(eta_row, W) - b)
    p_row = np.array([legendre_p(n, math.sin(lat[j + ng])) for j in range(ny)])
    w_row = np.array([rows["area"][j + ng] * nx for j in range(ny)])
    norm = 0.0
    for j in range(ny):
        norm += w_row[j] * p_row[j] * p_row[j]
    coef = (w_row / nx) * p_row
    return dict(nx=nx, ny=ny, rows=rows, lat=lat, b=b, h=h, coef=coef, norm=norm)


def projection(coef, b, h, nx, ny):
    """globe.mjs:1417-1424 on host arrays, in its order (rows left to right, then
    the rows in order); np.cumsum adds sequentially."""
    W = nx + 4
    eta = (b + h).reshape(ny + 4, W)[2:-2, 2:-2]
    rows = np.cumsum(eta, axis=1)[:, -1]
    return float(np.cumsum(coef * rows)[-1])


def crossings(times, vals, c0):
    """The zero crossings of globe.mjs:1432-1440: linear interpolation between
    the samples that straddle zero, from (t = 0, c0) on."""
    cross = []
    prev_t, prev_c = 0.0, c0
    for t, c in zip(times, vals):
        if (prev_c < 0) != (c < 0):
            cross.append(prev_t + (t - prev_t) * (0 - prev_c) / (c - prev_c))
        prev_t, prev_c = t, c
    return cross


def measure_mode_period(n=2, deg=6, H=4000.0, R=6371e3, g_ref=None, amp=1.0, periods=2, cfl=0.45,
                        lat0=-90.0, lat1=90.0, order=2, max_steps=50_000_000):
    """measureModePeriod (globe.mjs:1387-1456) on the GPU.  g_ref is required: a
    closed form built from the solver's own g rescales with any error in it."""
    if not (g_ref is not None and g_ref > 0):
        raise ValueError("measure_mode_period: g_ref is required; a closed form built from the solver's own g "
                         "can never detect an error in it")
    m = mode_problem(n, deg, H, R, amp, lat0, lat1)
    nx, ny = m["nx"], m["ny"]
    closed = 2 * math.pi * R / math.sqrt(g_ref * H * n * (n + 1))
    t_end = periods * closed
    s = swe_amd.Solver(nx, ny, m["rows"], sphere=True, R=R, manning=0.0, cfl=cfl, order=order)
    try:
        z = np.zeros_like(m["h"])
        s.load_state(m["b"], m["h"], z, z)
        s.set_bc(("periodic", "periodic", "reflect", "reflect"))
        # room for every step: t_end over the smallest step the grid's CFL allows a
        # resting ocean, with a margin for the mode's own speeds
        tax = polar_timestep_tax(m["rows"], m["lat"], cfl, H)
        cap = int(1.25 * t_end / tax["pole"]) + 64
        s.set_projection(m["coef"], cap)
        c0 = projection(m["coef"], m["b"], m["h"], nx, ny)
        s.autotune(reps=3)          # results do not depend on the chunk height
        s.use_graph(True)
        steps = s.advance(max_steps, swe_amd.DT_UNTIL, t_end=t_end)
        ms = s.last_advance_ms()
        times, vals = s.projection_series()
    finally:
        s.close()
    # the reference crosses zero on the normalised projection (s / norm)
    c0 = c0 / m["norm"]
    vals = vals / m["norm"]
    cross = crossings(times, vals, c0)
    base = dict(closed=closed, steps=steps, crossings=len(cross), amp0=c0,
                amp_end=float(vals[-1]) if len(vals) else c0, nx=nx, ny=ny, deg=deg, lat0=lat0, lat1=lat1,
                gpu_ms=ms, cell_updates=nx * ny * steps, polar_tax=tax)
    if len(cross) < 2:
        return dict(base, measured=math.nan, rel=math.nan, note="fewer than two zero crossings: nothing to measure")
    measured = 2 * (cross[-1] - cross[0]) / (len(cross) - 1)
    return dict(base, measured=measured, rel=(measured - closed) / closed)
