"""Seeded synthetic workloads of the benchmark configurations.

Every random number comes from SplitMix64 (Steele, Lea & Flood, 2014) seeded per
configuration, mapped to doubles in [0,1) by the top 53 bits; normals by
Box-Muller.  Inputs are built on the host with numpy and uploaded once.
"""
import math

import numpy as np

G_REF = 9.80665
M2_PERIOD_S = 12.42 * 3600.0


class SplitMix64:
    def __init__(self, seed):
        self.x = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.x = (self.x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def normals(self, n):
        """n standard normals (vectorised Box-Muller over SplitMix64 draws)."""
        m = (n + 1) // 2
        raw = np.empty(2 * m, dtype=np.uint64)
        # vectorised SplitMix64: states x + k*gamma for k = 1..2m
        gamma = np.uint64(0x9E3779B97F4A7C15)
        ks = np.arange(1, 2 * m + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.x) + ks * gamma
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        raw[:] = z
        self.x = (self.x + 2 * m * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        u = (raw >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        u1 = np.maximum(u[0::2], 1e-300)
        u2 = u[1::2]
        r = np.sqrt(-2.0 * np.log(u1))
        out = np.empty(2 * m)
        out[0::2] = r * np.cos(2 * np.pi * u2)
        out[1::2] = r * np.sin(2 * np.pi * u2)
        return out[:n]


def large_grid(nx=4096, ny=4096, dx=10.0, seed=2):
    """4096x4096 Cartesian, dx = 10 m: a plane beach from -30 m to +3 m along x, 64 Gaussian
    bumps (height U(-5,5) m, sigma U(50,400) m), N(0,0.2) m noise; eta0 = 0, u = v = 0.
    West Flather tide 1 m, period 12.42 h; Manning 0.025; f = 1e-4; CFL 0.45.
    Returns (b, h, hu, hv) over (ny+4)*(nx+4) cells and the run parameters."""
    rng = SplitMix64(seed)
    W, H = nx + 4, ny + 4
    x = (np.arange(W) - 2 + 0.5) * dx
    y = (np.arange(H) - 2 + 0.5) * dx
    Lx, Ly = nx * dx, ny * dx
    b = np.empty((H, W))
    b[:] = -30.0 + 33.0 * (x / Lx)
    for _ in range(64):
        cx, cy = rng.uniform(0, Lx), rng.uniform(0, Ly)
        amp, sig = rng.uniform(-5, 5), rng.uniform(50, 400)
        i0, i1 = np.searchsorted(x, cx - 5 * sig), np.searchsorted(x, cx + 5 * sig)
        j0, j1 = np.searchsorted(y, cy - 5 * sig), np.searchsorted(y, cy + 5 * sig)
        gx = np.exp(-((x[i0:i1] - cx) / sig) ** 2)
        gy = np.exp(-((y[j0:j1] - cy) / sig) ** 2)
        b[j0:j1, i0:i1] += amp * np.outer(gy, gx)
    b += 0.2 * rng.normals(H * W).reshape(H, W)
    h = np.maximum(0.0, 0.0 - b)
    z = np.zeros(H * W)
    params = dict(nx=nx, ny=ny, dx=dx, manning=0.025, coriolis=1e-4, cfl=0.45,
                  bc=("flather", "reflect", "reflect", "reflect"),
                  tide_amp=1.0, tide_omega=2 * math.pi / M2_PERIOD_S)
    return (b.ravel(), h.ravel(), z, z.copy()), params


def ensemble(nb=512, lo=64, hi=1024, dx=5.0, seed=3, with_state=True):
    """Batched ensemble of quarter-wave resonators: nb basins, nx and ny drawn
    log-uniform in [lo, hi]; each a closed rectangular channel of depth U(2,20) m
    with N(0,0.1) m roughness, Flather forcing at the west end (amplitude 0.1 m,
    period U(0.5,2) x 4L/sqrt(g h), incoming velocity sqrt(g/h) x eta), walls
    elsewhere, Manning 0.02, CFL 0.45.  Yields one dict per basin."""
    rng = SplitMix64(seed)
    out = []
    for i in range(nb):
        nx = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        ny = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        depth = rng.uniform(2.0, 20.0)
        L = nx * dx
        c = math.sqrt(G_REF * depth)
        period = rng.uniform(0.5, 2.0) * 4.0 * L / c
        spec = dict(nx=nx, ny=ny, dx=dx, depth=depth, period=period, manning=0.02, cfl=0.45,
                    bc=("flather", "reflect", "reflect", "reflect"), amp=0.1,
                    omega=2 * math.pi / period, uamp=0.1 * math.sqrt(G_REF / depth))
        if with_state:
            W, H = nx + 4, ny + 4
            b = -depth + 0.1 * rng.normals(W * H)
            h = np.maximum(0.0, -b)
            z = np.zeros(W * H)
            spec["state"] = (b, h, z, z.copy())
        out.append(spec)
    return out


def thacker_bowl(n=256, g=G_REF, A=0.5, h0=1.0, a=1000.0):
    """Thacker's (1981) radially symmetric oscillation in a paraboloid bowl:
    bed z = h0 (r/a)^2 on an n x n grid of the bowl's diameter / 200 (dx = 2a/200),
    centred; u = v = 0 and the free surface of the closed form at t = 0.
    Returns (state, params, exact) where exact(t) gives the closed-form depth."""
    dx = 2 * a / 200
    L = n * dx
    c = L / 2
    W = n + 4
    x = (np.arange(W) - 2 + 0.5) * dx
    X, Y = np.meshgrid(x, x)
    r2 = (X - c) ** 2 + (Y - c) ** 2
    omega = math.sqrt(8 * g * h0) / a
    period = 2 * math.pi / omega
    bed = h0 * r2 / (a * a)

    def surface(t):
        q = 1 - A * math.cos(omega * t)
        return h0 + h0 * (math.sqrt(1 - A * A) / q - 1 - (r2 / (a * a)) * ((1 - A * A) / (q * q) - 1))

    def exact_depth(t):
        return np.maximum(0.0, surface(t) - bed).ravel()

    h = exact_depth(0.0)
    z = np.zeros(W * W)
    params = dict(nx=n, ny=n, dx=dx, g=g, manning=0.0, period=period, bc=("reflect",) * 4)
    return (bed.ravel(), h, z, z.copy()), params, exact_depth


def inversion(nx=1024, ny=1024, dx=50.0, seed=4, ngauges=32):
    """Source inversion on a 1024x1024 Cartesian grid, dx = 50 m: an offshore shelf
    rising from -200 m (west) to -10 m, then a coast over the last 3 % of x that
    climbs to +5 m (wet/dry on the east edge); 32 Gaussian seamounts (height
    U(50,180) m, sigma U(1,4) km); N(0,0.5) m roughness.  Hidden truth: a sum of
    8 Gaussian humps of the surface, amplitude U(-1,1) m, sigma U(1,5) km, u = v = 0.
    Open (zero-gradient) west side, walls elsewhere; Manning 0.025; f = 1e-4; CFL 0.45.
    Gauges: 32 seeded wet cells at least 10 cells from the edges.
    Returns (b, h_truth, params, gauges)."""
    rng = SplitMix64(seed)
    W, H = nx + 4, ny + 4
    x = (np.arange(W) - 2 + 0.5) * dx
    y = (np.arange(H) - 2 + 0.5) * dx
    Lx, Ly = nx * dx, ny * dx
    xf = x / Lx
    shelf = np.where(xf < 0.97, -200.0 + 190.0 * np.clip(xf, 0, 0.97) / 0.97, -10.0 + 15.0 * (xf - 0.97) / 0.03)
    b = np.empty((H, W))
    b[:] = shelf

    def humps(n, amp_lo, amp_hi, s_lo, s_hi, into):
        for _ in range(n):
            cx, cy = rng.uniform(0, Lx), rng.uniform(0, Ly)
            amp, sig = rng.uniform(amp_lo, amp_hi), rng.uniform(s_lo, s_hi)
            into += amp * np.outer(np.exp(-((y - cy) / sig) ** 2), np.exp(-((x - cx) / sig) ** 2))

    humps(32, 50.0, 180.0, 1000.0, 4000.0, b)
    b += 0.5 * rng.normals(H * W).reshape(H, W)
    b = np.minimum(b, np.where(xf < 0.97, -2.0, np.inf)[None, :])   # seamounts stay submerged offshore
    eta = np.zeros((H, W))
    humps(8, -1.0, 1.0, 1000.0, 5000.0, eta)
    h = np.maximum(0.0, eta - b)
    gauges = []
    while len(gauges) < ngauges:
        i = 2 + 10 + int(rng.uniform(0, nx - 20))
        j = 2 + 10 + int(rng.uniform(0, ny - 20))
        k = j * W + i
        if h[j, i] > 5.0 and k not in gauges:
            gauges.append(k)
    params = dict(nx=nx, ny=ny, dx=dx, manning=0.025, coriolis=1e-4, cfl=0.45,
                  bc=("outflow", "reflect", "reflect", "reflect"), nsteps=4000, every=10)
    return b.ravel(), h.ravel(), params, np.array(gauges, dtype=np.int32)


def inversion_first_guess(b, h_truth, nx=1024, ny=1024, seed=45, amp=0.05, corr=3.0):
    """The first guess of the inversion's finite-difference checks: half the
    hidden truth surface plus `amp` metres of smooth seeded noise (white
    normals filtered by a Gaussian of `corr` cells, rescaled to unit deviation),
    over the wet cells; returns the depth h over W*H cells.  The noise keeps
    the surface differences away from zero, where every limiter sits on its
    kink, and being smooth it gives neighbouring differences one sign, so few
    limiter arguments lie near a kink either."""
    from scipy import ndimage
    W, H = nx + 4, ny + 4
    eta_t = np.where(h_truth > 0, b + h_truth, 0.0)
    n = ndimage.gaussian_filter(SplitMix64(seed).normals(W * H).reshape(H, W), corr).ravel()
    n /= n.std()
    return np.maximum(0.0, 0.5 * eta_t + amp * n - b)


def inversion_directions(n, nx, ny, dx, wet, seed=46, near=None, reach=2000.0):
    """n seeded random directions (dh, dn) of the inversion's control space, for
    finite-difference checks of its gradient: dh is 1 cm times a sum of 4 Gaussian
    humps of the surface (amplitude U(-1,1), sigma U(1,5) km, like the truth), dn is
    1e-3 times a low Fourier mode of the Manning field (1 to 6 waves per side, random
    phase); both vanish off the cells where wet (a bool array over W*H) is set.
    Hump centres are uniform over the domain, or, with `near` (cell indices k =
    j*W + i, e.g. the gauges), within `reach` metres of a randomly chosen one of
    those cells: inside the region a short run's gauges can see."""
    rng = SplitMix64(seed)
    W, H = nx + 4, ny + 4
    x = (np.arange(W) - 2 + 0.5) * dx
    y = (np.arange(H) - 2 + 0.5) * dx
    m = np.asarray(wet, dtype=bool).reshape(H, W)
    out = []
    for _ in range(n):
        dh = np.zeros((H, W))
        for _ in range(4):
            if near is None:
                cx, cy = rng.uniform(0, nx * dx), rng.uniform(0, ny * dx)
            else:
                k = int(near[min(len(near) - 1, int(rng.uniform(0, len(near))))])
                cx = (k % W - 2 + 0.5) * dx + rng.uniform(-reach, reach)
                cy = (k // W - 2 + 0.5) * dx + rng.uniform(-reach, reach)
            a, sg = rng.uniform(-1, 1), rng.uniform(1000, 5000)
            dh += 0.01 * a * np.outer(np.exp(-((y - cy) / sg) ** 2), np.exp(-((x - cx) / sg) ** 2))
        kx, ky, ph = rng.uniform(1, 6), rng.uniform(1, 6), rng.uniform(0, 2 * math.pi)
        dn = 1e-3 * np.outer(np.cos(2 * math.pi * ky * y / (ny * dx)), np.sin(2 * math.pi * kx * x / (nx * dx) + ph))
        out.append(((dh * m).ravel(), (dn * m).ravel()))
    return out


def sphere_tides(nx=2160, ny=1080, members=16, seed=5):
    """Global ocean on the 10-arc-minute lat-lon sphere (2160x1080, R = 6371 km,
    Omega = 7.292115e-5 s^-1, the globe's metric rows): bathymetry from spectral
    noise with k^-2 amplitude (seeded normals shaped in Fourier space), scaled
    linearly to -6000 .. +2000 m, land dry; an ocean at rest (eta0 = 0).
    Equilibrium tide M2 0.5 m, S2 0.2 m, K1 0.3 m.  Members differ only in their
    Manning n, U(0.015, 0.035).  Returns (b, h, params, manning per member)."""
    rng = SplitMix64(seed)
    W, H = nx + 4, ny + 4
    wn = rng.normals(ny * nx).reshape(ny, nx)
    F = np.fft.rfft2(wn)
    ky = np.fft.fftfreq(ny)[:, None] * ny
    kx = np.fft.rfftfreq(nx)[None, :] * nx
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    F *= k ** -2.0
    F[0, 0] = 0.0
    z = np.fft.irfft2(F, s=(ny, nx))
    z = -6000.0 + 8000.0 * (z - z.min()) / (z.max() - z.min())
    b = np.empty((H, W))
    b[2:-2, 2:-2] = z
    b[2:-2, :2] = z[:, -2:]                  # periodic seam
    b[2:-2, -2:] = z[:, :2]
    b[:2] = b[3:1:-1]                        # mirrored beyond the poles
    b[-2:] = b[-3:-5:-1]
    h = np.maximum(0.0, -b)
    manning = [rng.uniform(0.015, 0.035) for _ in range(members)]
    params = dict(nx=nx, ny=ny, R=6371e3, omega=7.292115e-5, lat0=-90.0, lat1=90.0, cfl=0.45,
                  tides=[(2, 0.5, 1.405189025e-4, 0.0), (2, 0.2, 1.454441043e-4, 0.0), (1, 0.3, 7.292115e-5, 0.0)],
                  days=5.0)
    return b.ravel(), h.ravel(), params, manning
