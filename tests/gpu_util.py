"""Helpers for GPU tests: seeded synthetic basins built the same way for the
HIP solver and for the CPU oracle."""
import math

import numpy as np

import oracle
import swe_amd


def splitmix64(seed):
    """SplitMix64 (Steele, Lea & Flood 2014) as a generator of uint64."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def random_basin(nx, ny, seed, *, dx=5.0, wetdry=True):
    """Initial b, h, hu, hv (full W*H arrays) of a rough basin with a beach and a bump."""
    rng = np.random.default_rng(seed)
    W, H = nx + 4, ny + 4
    x = (np.arange(W) - 2 + 0.5) * dx
    y = (np.arange(H) - 2 + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    L = nx * dx
    b = -3.0 + (4.0 * X / L if wetdry else 0.0) + 0.3 * np.sin(Y / (7 * dx)) + rng.normal(0, 0.05, X.shape)
    eta = 0.4 * np.exp(-(((X - 0.3 * L) ** 2 + (Y - 0.5 * ny * dx) ** 2) / (0.02 * L * L)))
    h = np.maximum(0.0, eta - b)
    hu = h * rng.normal(0, 0.2, X.shape)
    hv = h * rng.normal(0, 0.2, X.shape)
    hu[h <= 1e-3] = 0
    hv[h <= 1e-3] = 0
    return b.ravel(), h.ravel(), hu.ravel(), hv.ravel()


def pair(nx, ny, seed, *, bc=("reflect",) * 4, dx=5.0, manning=0.0, coriolis=0.0, order=2,
         limiter="mc", manning_field=False, sponge=False, wetdry=True):
    """(gpu Solver, OracleSim) loaded with the same seeded basin."""
    rows = swe_amd.geometry_cartesian(nx, ny, dx, coriolis=coriolis)
    kw = dict(order=order, limiter=limiter, manning=manning, coriolis=coriolis, dx=dx, dy=dx)
    g = swe_amd.Solver(nx, ny, rows, **kw)
    o = oracle.OracleSim(nx, ny, rows, use_manning_field=manning_field, **kw)
    st = random_basin(nx, ny, seed, dx=dx, wetdry=wetdry)
    g.load_state(*st)
    o.load_state(*st)
    g.set_bc(bc)
    o.set_bc(bc)
    if manning_field:
        n = np.random.default_rng(seed + 1).uniform(0.0, 0.05, (ny + 4) * (nx + 4))
        n[::7] = 0.0
        g.set_manning_field(n)
        o.set_manning_field(n)
    if sponge:
        s = np.zeros((ny + 4, nx + 4))
        s[2:-2, -10:-2] = 0.05 * np.linspace(0, 1, 8)
        g.set_sponge(s.ravel())
        o.set_sponge(s.ravel())
    return g, o


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def first_diff(a, b, W):
    d = np.nonzero(bits(a) != bits(b))[0]
    if d.size == 0:
        return None
    k = int(d[0])
    return (k % W, k // W, float(a[k]), float(b[k]), int(d.size))
