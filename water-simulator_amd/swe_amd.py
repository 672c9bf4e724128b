"""Python host mirror of the HIP shallow-water step (ctypes over include/swe_amd.h).

`Solver` mirrors ShallowWater of src/swe.mjs (the reference solver): the same
state arrays (b, h, hu, hv over W*H cells with two ghost cells per side), the
same boundary kinds, step(dt) with dt=None meaning self-timed, maxDt(),
applyBC(), residual(dt), setManningField() and the sponge forcing.  All
arithmetic runs in libswe_amd.so on the GPU; there is no CPU fallback: if the
library or a GPU is missing, construction raises.

`geometry_cartesian` / `geometry_spherical` restate src/geometry.mjs (the
metric rows the step divides by).
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SWE_AMD_LIB") or os.path.join(HERE, "libswe_amd.so")

NROWS = 9
ROWS = ["dxRow", "dyRowN", "dyRowS", "dyCFL", "bedPhi", "tanPhi", "fRow", "geoCoef", "area"]
F_B, F_H, F_HU, F_HV, F_MANNING, F_RH, F_RHU, F_RHV, F_THETA, F_SPONGE = range(10)
BC = {"reflect": 0, "outflow": 1, "periodic": 2, "flather": 3, "stored": 4}
LIMITER = {"minmod": 0, "mc": 1, "central": 2}
(S_MANNING, S_CORIOLIS, S_CFL, S_T, S_SPONGE_REF, S_STEPS, S_MASS_FLOORED, S_DT_NEXT, S_NONFINITE,
 S_CHUNK_ROWS, S_DIAG_MS) = range(11)
DT_SELF, DT_FIXED, DT_UNTIL, DT_SAMPLED = 0, 1, 2, 3
DIAGNOSTICS = ("volume", "energy", "maxSurfaceDeviation", "maxSpeed", "finite")


class SweError(RuntimeError):
    pass


class Desc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("order", ctypes.c_int32), ("limiter", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("g", ctypes.c_double), ("min_depth", ctypes.c_double), ("eps4", ctypes.c_double),
                ("cfl", ctypes.c_double), ("coriolis", ctypes.c_double), ("manning", ctypes.c_double),
                ("R", ctypes.c_double), ("rows", ctypes.POINTER(ctypes.c_double) * NROWS)]


_lib = None


def lib():
    """Load libswe_amd.so (built by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SweError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        sig = {
            "swe_create": (i, [ctypes.POINTER(Desc), ctypes.POINTER(vp)]),
            "swe_destroy": (None, [vp]),
            "swe_last_error": (ctypes.c_char_p, [vp]),
            "swe_build_info": (ctypes.c_char_p, []),
            "swe_js_pow": (d, [d, d]),
            "swe_set_field": (i, [vp, i, dp]),
            "swe_get_field": (i, [vp, i, dp]),
            "swe_set_row": (i, [vp, i, dp]),
            "swe_set_bc": (i, [vp, i, i]),
            "swe_set_scalar": (i, [vp, i, d]),
            "swe_get_scalar": (i, [vp, i, dp]),
            "swe_apply_bc": (i, [vp, dp]),
            "swe_residual": (i, [vp, d]),
            "swe_max_dt": (i, [vp, dp]),
            "swe_step": (i, [vp, d, dp, i, i, dp, dp]),
            "swe_advance": (i, [vp, i64, i, d, d, ctypes.POINTER(i64)]),
            "swe_set_harmonic": (i, [vp, i, i, d, dp, dp, dp, i, d, dp, dp, dp]),
            "swe_set_harmonic_ramp": (i, [vp, i, d, i, d]),
            "swe_use_graph": (i, [vp, i]),
            "swe_set_chunk_rows": (i, [vp, i]),
            "swe_sync": (i, [vp]),
            "swe_device_ptr": (i, [vp, i, ctypes.POINTER(dp)]),
            "swe_last_advance_ms": (i, [vp, dp]),
            "swe_profile_step": (i, [vp, i, dp]),
            "swe_autotune": (i, [vp, i, ctypes.POINTER(i)]),
            "swe_set_tide": (i, [vp, i, ctypes.POINTER(ctypes.c_int32), dp, dp, dp, dp, dp]),
            "swe_diagnostics": (i, [vp, d, dp]),
            "swe_set_recorder": (i, [vp, i, ctypes.POINTER(ctypes.c_int32), i, d, d, i64]),
            "swe_get_record": (i, [vp, dp, dp, ctypes.POINTER(i64)]),
            "swe_set_projection": (i, [vp, dp, i64]),
            "swe_get_projection": (i, [vp, dp, dp, ctypes.POINTER(i64)]),
            "swe_step_begin": (i, [vp, d, dp, dp]),
            "swe_step_end": (i, [vp, dp, i, i, dp]),
            "swe_keep_residual": (i, [vp, i]),
            "swe_get_ext": (i, [vp, dp]),
            "swe_get_ghosts": (i, [vp, i, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ext(ext):
    e = np.zeros(8)
    if ext is not None:
        e[:] = ext
    return e


def js_pow4(x):
    """max(minDepth,1e-12)**4 as the JS engine evaluates it (the library's fdlibm pow)."""
    return lib().swe_js_pow(x, 4.0)


class Solver:
    """One basin on the GPU.  Mirrors the reference ShallowWater's state and methods."""

    def __init__(self, nx, ny, rows, *, sphere=False, order=2, limiter="mc", g=9.80665, min_depth=1e-3,
                 eps4=None, cfl=0.45, coriolis=0.0, manning=0.025, R=math.inf, dx=None, dy=None,
                 use_manning_field=False):
        L = lib()
        if eps4 is None:
            eps4 = js_pow4(max(min_depth, 1e-12))
        self.nx, self.ny, self.W, self.H = nx, ny, nx + 4, ny + 4
        self.N = self.W * self.H
        self.sphere = sphere
        self.min_depth = min_depth
        self.dx, self.dy = dx, dy
        self._rows = {r: np.ascontiguousarray(rows[r], dtype=np.float64) for r in ROWS}
        d = Desc()
        d.nx, d.ny, d.kind, d.order = nx, ny, int(sphere), order
        d.limiter = LIMITER[limiter] if isinstance(limiter, str) else limiter
        d.g, d.min_depth, d.eps4, d.cfl = g, min_depth, eps4, cfl
        d.coriolis, d.manning, d.R = coriolis, manning, R
        for k, r in enumerate(ROWS):
            d.rows[k] = _dp(self._rows[r])
        h = ctypes.c_void_p()
        rc = L.swe_create(ctypes.byref(d), ctypes.byref(h))
        if rc != 0:
            raise SweError(f"swe_create: {L.swe_last_error(None).decode()}")
        self._h = h
        self._mf = 0.0
        self._probe_rh = False
        self._rec_n = 0

    def _ck(self, rc, what):
        if rc != 0:
            raise SweError(f"{what}: {lib().swe_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            lib().swe_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state --------------------------------------------------------------
    def set_field(self, f, a):
        if a is None:
            self._ck(lib().swe_set_field(self._h, f, None), "swe_set_field")
            return
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.N
        self._ck(lib().swe_set_field(self._h, f, _dp(a)), "swe_set_field")

    def get_field(self, f):
        out = np.empty(self.N)
        self._ck(lib().swe_get_field(self._h, f, _dp(out)), "swe_get_field")
        return out

    field = get_field

    def load_state(self, b, h, hu, hv):
        for f, a in ((F_B, b), (F_H, h), (F_HU, hu), (F_HV, hv)):
            self.set_field(f, a)

    def state(self):
        return tuple(self.get_field(f) for f in (F_B, F_H, F_HU, F_HV))

    @property
    def h(self):
        return self.get_field(F_H)

    @property
    def hu(self):
        return self.get_field(F_HU)

    @property
    def hv(self):
        return self.get_field(F_HV)

    @property
    def b(self):
        return self.get_field(F_B)

    def set_manning_field(self, n):
        self.set_field(F_MANNING, n)

    def set_sponge(self, strength, eta_ref=0.0):
        self.set_field(F_SPONGE, strength)
        self.set_scalar(S_SPONGE_REF, eta_ref)

    def set_row(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._ck(lib().swe_set_row(self._h, ROWS.index(which) if isinstance(which, str) else which, _dp(a)),
                 "swe_set_row")

    def set_bc(self, kinds):
        for side, k in enumerate(kinds):
            self._ck(lib().swe_set_bc(self._h, side, BC[k] if isinstance(k, str) else k), "swe_set_bc")

    def set_scalar(self, which, v):
        self._ck(lib().swe_set_scalar(self._h, which, float(v)), "swe_set_scalar")

    def get_scalar(self, which):
        v = ctypes.c_double()
        self._ck(lib().swe_get_scalar(self._h, which, ctypes.byref(v)), "swe_get_scalar")
        return v.value

    @property
    def t(self):
        return self.get_scalar(S_T)

    def set_time(self, t, steps=0):
        self.set_scalar(S_T, t)
        self.set_scalar(S_STEPS, steps)

    # ---- the step ---------------------------------------------------------------
    def apply_bc(self, ext=None):
        e = _ext(ext)
        self._ck(lib().swe_apply_bc(self._h, _dp(e)), "swe_apply_bc")

    def residual(self, dt):
        self._ck(lib().swe_residual(self._h, float(dt)), "swe_residual")

    def max_dt(self):
        v = ctypes.c_double()
        self._ck(lib().swe_max_dt(self._h, ctypes.byref(v)), "swe_max_dt")
        return v.value

    def step(self, dt=None, ext=None, want_theta=False, want_rh=False):
        e = _ext(ext)
        taken, mf = ctypes.c_double(), ctypes.c_double()
        self._ck(lib().swe_step(self._h, math.nan if dt is None else float(dt), _dp(e), int(want_theta),
                                int(want_rh), ctypes.byref(taken), ctypes.byref(mf)), "swe_step")
        self._mf = mf.value
        return taken.value

    def step_begin(self, dt=None, ext=None):
        """First half of step(): dt and stage 1; the fields then hold the stage-1
        state until step_end() (for a host ghost fill in between).  Returns dt (0:
        refused, no step_end follows)."""
        e = _ext(ext)
        taken = ctypes.c_double()
        self._ck(lib().swe_step_begin(self._h, math.nan if dt is None else float(dt), _dp(e), ctypes.byref(taken)),
                 "swe_step_begin")
        return taken.value

    def step_end(self, ext=None, want_theta=False, want_rh=False):
        """Second half of step(): the second applyBC, stage 2 and the epilogue."""
        e = None if ext is None else _ext(ext)
        mf = ctypes.c_double()
        self._ck(lib().swe_step_end(self._h, None if e is None else _dp(e), int(want_theta), int(want_rh),
                                    ctypes.byref(mf)), "swe_step_end")
        self._mf = mf.value
        return mf.value

    def keep_residual(self, on=True):
        """Every later step also leaves stage 2's theta and rh/rhu/rhv in the probe fields."""
        self._ck(lib().swe_keep_residual(self._h, int(on)), "swe_keep_residual")

    def ext(self):
        """The Flather (eta_ext, u_ext) per side the next step uses (device-evaluated in advance())."""
        out = np.zeros(8)
        self._ck(lib().swe_get_ext(self._h, _dp(out)), "swe_get_ext")
        return out

    def get_ghosts(self, f, into):
        """Copy the ghost ring of field f into the W*H array `into` (interior untouched)."""
        assert into.dtype == np.float64 and into.size == self.N and into.flags.c_contiguous
        self._ck(lib().swe_get_ghosts(self._h, f, _dp(into)), "swe_get_ghosts")

    def sponge(self, dt, ref):
        """The sponge forcing runs inside step() on the device (set_sponge)."""

    def advance(self, nsteps, dt_mode=DT_SELF, dt_value=0.0, t_end=math.inf):
        n = ctypes.c_int64()
        self._ck(lib().swe_advance(self._h, int(nsteps), dt_mode, float(dt_value), float(t_end), ctypes.byref(n)),
                 "swe_advance")
        return n.value

    def last_advance_ms(self):
        v = ctypes.c_double()
        self._ck(lib().swe_last_advance_ms(self._h, ctypes.byref(v)), "swe_last_advance_ms")
        return v.value

    # ---- device-resident diagnostics and gauges ----------------------------------
    def diagnostics(self, eta_ref=0.0):
        """volume, energy, maxSurfaceDeviation(eta_ref), maxSpeed and finite of the
        device state (swe.mjs:819-917) without copying it to the host."""
        out = np.zeros(5)
        self._ck(lib().swe_diagnostics(self._h, float(eta_ref), _dp(out)), "swe_diagnostics")
        r = dict(zip(DIAGNOSTICS, out.tolist()))
        r["finite"] = bool(r["finite"])
        return r

    # the reference's method names over the device diagnostics
    def volume(self):
        return self.diagnostics()["volume"]

    def energy(self):
        return self.diagnostics()["energy"]

    def max_speed(self):
        return self.diagnostics()["maxSpeed"]

    def max_surface_deviation(self, ref):
        return self.diagnostics(ref)["maxSurfaceDeviation"]

    def set_recorder(self, cells, every=0, t0=0.0, dt_rec=0.0, capacity=1024):
        """Arm the gauge recorder of advance(): eta = b + h at `cells` after every
        `every`-th step, or (every=0) at the sample times t0, t0 + dt_rec, ...
        (tools/waves.mjs:211-232; advance(n, DT_SAMPLED, t_end=T) lands on them).
        An empty `cells` disarms it."""
        c = np.ascontiguousarray(cells, dtype=np.int32)
        self._rec_n = c.size
        self._ck(lib().swe_set_recorder(self._h, c.size, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                        int(every), float(t0), float(dt_rec), int(capacity)), "swe_set_recorder")

    def set_projection(self, coef, capacity):
        """Arm the projection recorder of advance(): after every step that runs,
        sum_j coef[j] * (row sum of b + h) in the reference's order (the mode
        projection of globe.mjs:1417-1424).  coef=None disarms."""
        if coef is None:
            self._ck(lib().swe_set_projection(self._h, None, 0), "swe_set_projection")
            return
        c = np.ascontiguousarray(coef, dtype=np.float64)
        if c.size != self.ny:
            raise ValueError(f"projection: {c.size} coefficients for {self.ny} rows")
        self._ck(lib().swe_set_projection(self._h, _dp(c), int(capacity)), "swe_set_projection")

    def projection_series(self):
        """(times, values) of the projection recorder so far."""
        n = ctypes.c_int64()
        self._ck(lib().swe_get_projection(self._h, None, None, ctypes.byref(n)), "swe_get_projection")
        vals, times = np.zeros(n.value), np.zeros(n.value)
        self._ck(lib().swe_get_projection(self._h, _dp(vals), _dp(times), ctypes.byref(n)), "swe_get_projection")
        return times, vals

    def record(self):
        """(times, eta): the samples so far, eta of shape (samples, gauges)."""
        n = ctypes.c_int64()
        self._ck(lib().swe_get_record(self._h, None, None, ctypes.byref(n)), "swe_get_record")
        eta = np.zeros((n.value, self._rec_n))
        times = np.zeros(n.value)
        self._ck(lib().swe_get_record(self._h, _dp(eta), _dp(times), ctypes.byref(n)), "swe_get_record")
        return times, eta

    def set_harmonic(self, side, amp=(), omega=(), phase=(), msl=0.0, uamp=(), uomega=(), uphase=(), u0=0.0,
                     ramp=0.0, u_scale=None):
        """Flather forcing of swe_advance: eta = msl + sum amp cos(omega t - phase),
        u likewise; ramp > 0 fades the amplitudes in over `ramp` seconds and
        u_scale makes u = u_scale (eta - msl) (regularWave, src/waves.mjs:163-175)."""
        def arr(x):
            a = np.zeros(4)
            a[:len(x)] = x
            return a
        A, O, P, UA, UO, UP = (arr(x) for x in (amp, omega, phase, uamp, uomega, uphase))
        self._keep = (A, O, P, UA, UO, UP)
        self._ck(lib().swe_set_harmonic(self._h, side, len(amp), msl, _dp(A), _dp(O), _dp(P), len(uamp), u0,
                                        _dp(UA), _dp(UO), _dp(UP)), "swe_set_harmonic")
        if ramp > 0 or u_scale is not None:
            self._ck(lib().swe_set_harmonic_ramp(self._h, side, float(ramp), int(u_scale is not None),
                                                 float(u_scale or 0.0)), "swe_set_harmonic_ramp")

    def set_tide(self, constituents, lat, lon):
        """Equilibrium-tide body force (spherical grids): constituents = [(m, amp, omega, phase), ...],
        m = 2 semidiurnal, 1 diurnal; lat (H) and lon (W) cell-centre angles in radians."""
        self._ck(lib().swe_set_tide(self._h, *_tide_args(constituents, lat, lon)), "swe_set_tide")

    def profile_step(self, nsteps):
        """Mean kernel durations (ms) [stage 1, stage 2, finalize] over nsteps device steps."""
        out = np.zeros(3)
        self._ck(lib().swe_profile_step(self._h, int(nsteps), _dp(out)), "swe_profile_step")
        return out

    def chunk_rows(self):
        return int(self.get_scalar(S_CHUNK_ROWS))

    def autotune(self, reps=3):
        """Pick the fastest row-chunk height on the device; returns it."""
        best = ctypes.c_int()
        self._ck(lib().swe_autotune(self._h, int(reps), ctypes.byref(best)), "swe_autotune")
        return best.value

    def use_graph(self, on=True):
        self._ck(lib().swe_use_graph(self._h, int(on)), "swe_use_graph")

    def set_chunk_rows(self, rows):
        self._ck(lib().swe_set_chunk_rows(self._h, int(rows)), "swe_set_chunk_rows")

    def sync(self):
        self._ck(lib().swe_sync(self._h), "swe_sync")

    # ---- diagnostics over the host mirror (swe.mjs:819-917) ---------------------
    @property
    def mass_floored(self):
        return self._mf

    def finite(self):
        _, h, hu, hv = self.state()
        return bool(np.all(np.isfinite(h)) and np.all(np.isfinite(hu)) and np.all(np.isfinite(hv)))

    def interior(self, a):
        return a.reshape(self.H, self.W)[2:-2, 2:-2]


# ---- geometry (src/geometry.mjs) -----------------------------------------------
R_EARTH = 6371e3
OMEGA_EARTH = 7.292115e-5


# equilibrium-tide constituents: (m, angular frequency rad/s)
TIDE_M2 = (2, 1.405189025e-4)
TIDE_S2 = (2, 1.454441043e-4)
TIDE_K1 = (1, 7.292115e-5)


def _tide_args(constituents, lat, lon):
    n = len(constituents)
    m = (ctypes.c_int32 * max(n, 1))(*[int(c[0]) for c in constituents])
    keep = [np.ascontiguousarray([float(c[k]) for c in constituents] or [0.0]) for k in (1, 2, 3)]
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    _tide_args.keep = (keep, lat, lon)
    return n, m, _dp(keep[0]), _dp(keep[1]), _dp(keep[2]), _dp(lat), _dp(lon)


def sphere_lonlat(nx, ny, lat0=-90.0, lat1=90.0, ng=2):
    """Cell-centre latitudes (H rows) and longitudes (W columns) in radians, as
    cellLonLat (swe.mjs:262-272, geometry.mjs:314) places them: lon = (i + 0.5) dlam - pi,
    lat = the centre of the row's (clamped) faces."""
    H, W = ny + 2 * ng, nx + 2 * ng
    dlam = 2 * math.pi / nx
    p0, p1 = lat0 * math.pi / 180, lat1 * math.pi / 180
    dphi = (p1 - p0) / ny
    half_pi = math.pi / 2
    lat = np.empty(H)
    for j in range(H):
        jj = j - ng
        pS = p0 if jj == 0 else p1 if jj == ny else p0 + jj * dphi
        pN = p0 if jj + 1 == 0 else p1 if jj + 1 == ny else p0 + (jj + 1) * dphi
        pS = min(half_pi, max(-half_pi, pS))
        pN = min(half_pi, max(-half_pi, pN))
        lat[j] = 0.5 * (pS + pN)
    lon = (np.arange(W) - ng + 0.5) * dlam - math.pi
    return lat, lon


def geometry_cartesian(nx, ny, dx, dy=None, coriolis=0.0, ng=2):
    """cartesianGeometry (geometry.mjs:163-176): every row ASSIGNED the scalars."""
    dy = dx if dy is None else dy
    H = ny + 2 * ng
    rows = {r: np.zeros(H) for r in ROWS}
    rows["dxRow"][:] = dx
    rows["dyRowN"][:] = dy
    rows["dyRowS"][:] = dy
    rows["dyCFL"][:] = dy
    rows["bedPhi"][:] = dy
    rows["area"][:] = dx * dy
    rows["fRow"][:] = coriolis
    return rows


def _cos_lat(p):
    co = math.pi / 2 - abs(p)
    return 0.0 if co <= 0 else math.sin(co)


def geometry_spherical(nx, ny, R=R_EARTH, lat0=-90.0, lat1=90.0, omega=0.0, ng=2):
    """sphericalGeometry (geometry.mjs:190-299): product-form area, divisors as divisors,
    Infinity at a pole face, reflected latitudes for rows beyond a pole.  Transcendentals
    come from the C library, so rows can differ from the JS engine's in the last ulp."""
    if not lat1 > lat0:
        raise ValueError(f"lat1 ({lat1}) must exceed lat0 ({lat0})")
    if lat0 < -90 or lat1 > 90:
        raise ValueError(f"latitudes must lie in [-90, 90]; got {lat0}..{lat1}")
    touches = lat0 <= -90 + 1e-12 or lat1 >= 90 - 1e-12
    if touches and nx % 2 != 0:
        raise ValueError(f"nx must be even when the domain reaches a pole (got {nx})")
    H = ny + 2 * ng
    rows = {r: np.zeros(H) for r in ROWS}
    dlam = 2 * math.pi / nx
    p0, p1 = lat0 * math.pi / 180, lat1 * math.pi / 180
    dphi = (p1 - p0) / ny
    half_pi = math.pi / 2

    def face_lat(k):
        return p0 if k == 0 else p1 if k == ny else p0 + k * dphi

    def clamp(p):
        return half_pi if p > half_pi else -half_pi if p < -half_pi else p

    def reflect(p):
        return math.pi - p if p > half_pi else -math.pi - p if p < -half_pi else p

    dy = R * dphi
    for j in range(H):
        jj = j - ng
        rawS, rawN = face_lat(jj), face_lat(jj + 1)
        pS, pN = clamp(rawS), clamp(rawN)
        pC = 0.5 * (pS + pN)
        half = 0.5 * (pN - pS)
        lyS = R * dlam * _cos_lat(pS)
        lyN = R * dlam * _cos_lat(pN)
        lx = R * dphi
        A = 2 * R * R * dlam * _cos_lat(pC) * math.sin(half)
        if not A > 0:
            rS, rN = reflect(rawS), reflect(rawN)
            A = abs(2 * R * R * dlam * _cos_lat(0.5 * (rS + rN)) * math.sin(0.5 * (rN - rS)))
        rows["area"][j] = A
        rows["dxRow"][j] = A / lx
        rows["dyRowN"][j] = A / lyN if lyN > 0 else math.inf
        rows["dyRowS"][j] = A / lyS if lyS > 0 else math.inf
        rows["dyCFL"][j] = min(rows["dyRowN"][j], rows["dyRowS"][j], dy)
        rows["bedPhi"][j] = 2 * A / (lyN + lyS) if (lyN + lyS) > 0 else math.inf
        rows["tanPhi"][j] = math.tan(pC)
        rows["fRow"][j] = 2 * omega * math.sin(pC)
        rows["geoCoef"][j] = (lyN - lyS) / (4 * A)
    return rows


class Batch:
    """Many independent basins advanced together (swe_batch_* of the C ABI).

    basins: list of dicts with keys nx, ny, rows (geometry rows) and the Solver
    keyword arguments (order, limiter, g, min_depth, eps4, cfl, coriolis, manning,
    sphere, R).  Each basin keeps its own dt, t, boundaries and forcing."""

    def __init__(self, basins):
        L = lib()
        self._bind(L)
        n = len(basins)
        arr = (Desc * n)()
        self._keep = []
        self.shapes = []
        for i, bd in enumerate(basins):
            d = arr[i]
            d.nx, d.ny = bd["nx"], bd["ny"]
            d.kind = int(bd.get("sphere", False))
            d.order = bd.get("order", 2)
            lim = bd.get("limiter", "mc")
            d.limiter = LIMITER[lim] if isinstance(lim, str) else lim
            d.g = bd.get("g", 9.80665)
            d.min_depth = bd.get("min_depth", 1e-3)
            d.eps4 = bd.get("eps4") or js_pow4(max(d.min_depth, 1e-12))
            d.cfl = bd.get("cfl", 0.45)
            d.coriolis = bd.get("coriolis", 0.0)
            d.manning = bd.get("manning", 0.025)
            d.R = bd.get("R", math.inf)
            rows = {r: np.ascontiguousarray(bd["rows"][r], dtype=np.float64) for r in ROWS}
            self._keep.append(rows)
            for k, r in enumerate(ROWS):
                d.rows[k] = _dp(rows[r])
            self.shapes.append((bd["nx"] + 4, bd["ny"] + 4))
        h = ctypes.c_void_p()
        rc = L.swe_batch_create(arr, n, ctypes.byref(h))
        if rc != 0:
            raise SweError(f"swe_batch_create: {L.swe_batch_last_error(None).decode()}")
        self._h = h
        self.n = n

    @staticmethod
    def _bind(L):
        if getattr(L, "_batch_bound", False):
            return
        vp, i, d, dp, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
        sig = {
            "swe_batch_create": (i, [ctypes.POINTER(Desc), i, ctypes.POINTER(vp)]),
            "swe_batch_destroy": (None, [vp]),
            "swe_batch_last_error": (ctypes.c_char_p, [vp]),
            "swe_batch_set_field": (i, [vp, i, i, dp]),
            "swe_batch_get_field": (i, [vp, i, i, dp]),
            "swe_batch_set_bc": (i, [vp, i, i, i]),
            "swe_batch_set_harmonic": (i, [vp, i, i, i, d, dp, dp, dp, i, d, dp, dp, dp]),
            "swe_batch_get_scalar": (i, [vp, i, i, dp]),
            "swe_batch_get_ext": (i, [vp, i, dp]),
            "swe_batch_set_chunk_rows": (i, [vp, i, i]),
            "swe_batch_use_graph": (i, [vp, i]),
            "swe_batch_advance": (i, [vp, i64, ctypes.POINTER(i64)]),
            "swe_batch_last_advance_ms": (i, [vp, dp]),
            "swe_batch_autotune": (i, [vp, i, ctypes.POINTER(i), ctypes.POINTER(i)]),
            "swe_batch_set_tide": (i, [vp, i, i, ctypes.POINTER(ctypes.c_int32), dp, dp, dp, dp, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._batch_bound = True

    def _ck(self, rc, what):
        if rc != 0:
            raise SweError(f"{what}: {lib().swe_batch_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            lib().swe_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_field(self, basin, f, a):
        W, H = self.shapes[basin]
        if a is None:
            self._ck(lib().swe_batch_set_field(self._h, basin, f, None), "swe_batch_set_field")
            return
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == W * H
        self._ck(lib().swe_batch_set_field(self._h, basin, f, _dp(a)), "swe_batch_set_field")

    def get_field(self, basin, f):
        W, H = self.shapes[basin]
        out = np.empty(W * H)
        self._ck(lib().swe_batch_get_field(self._h, basin, f, _dp(out)), "swe_batch_get_field")
        return out

    def load_state(self, basin, b, h, hu, hv):
        for f, a in ((F_B, b), (F_H, h), (F_HU, hu), (F_HV, hv)):
            self.set_field(basin, f, a)

    def state(self, basin):
        return tuple(self.get_field(basin, f) for f in (F_B, F_H, F_HU, F_HV))

    def set_bc(self, basin, kinds):
        for side, k in enumerate(kinds):
            self._ck(lib().swe_batch_set_bc(self._h, basin, side, BC[k] if isinstance(k, str) else k),
                     "swe_batch_set_bc")

    def set_harmonic(self, basin, side, amp=(), omega=(), phase=(), msl=0.0, uamp=(), uomega=(), uphase=(), u0=0.0):
        def arr(x):
            a = np.zeros(4)
            a[:len(x)] = x
            return a
        A, O, P, UA, UO, UP = (arr(x) for x in (amp, omega, phase, uamp, uomega, uphase))
        self._ck(lib().swe_batch_set_harmonic(self._h, basin, side, len(amp), msl, _dp(A), _dp(O), _dp(P),
                                              len(uamp), u0, _dp(UA), _dp(UO), _dp(UP)), "swe_batch_set_harmonic")

    def set_tide(self, basin, constituents, lat, lon):
        self._ck(lib().swe_batch_set_tide(self._h, basin, *_tide_args(constituents, lat, lon)), "swe_batch_set_tide")

    def get_scalar(self, basin, which):
        v = ctypes.c_double()
        self._ck(lib().swe_batch_get_scalar(self._h, basin, which, ctypes.byref(v)), "swe_batch_get_scalar")
        return v.value

    def ext(self, basin):
        """The Flather (eta_ext, u_ext) per side basin `basin`'s next step uses."""
        out = np.zeros(8)
        self._ck(lib().swe_batch_get_ext(self._h, basin, _dp(out)), "swe_batch_get_ext")
        return out

    def set_chunk_rows(self, rows, nwaves=0):
        self._ck(lib().swe_batch_set_chunk_rows(self._h, int(rows), int(nwaves)), "swe_batch_set_chunk_rows")

    def use_graph(self, on=True):
        self._ck(lib().swe_batch_use_graph(self._h, int(on)), "swe_batch_use_graph")

    def advance(self, nsteps):
        m = ctypes.c_int64()
        self._ck(lib().swe_batch_advance(self._h, int(nsteps), ctypes.byref(m)), "swe_batch_advance")
        return m.value

    def last_advance_ms(self):
        v = ctypes.c_double()
        self._ck(lib().swe_batch_last_advance_ms(self._h, ctypes.byref(v)), "swe_batch_last_advance_ms")
        return v.value

    def autotune(self, steps=2):
        r, w = ctypes.c_int(), ctypes.c_int()
        self._ck(lib().swe_batch_autotune(self._h, int(steps), ctypes.byref(r), ctypes.byref(w)), "swe_batch_autotune")
        return r.value, w.value


# ---- discrete adjoint ----------------------------------------------------------

def _bind_adjoint(L):
    if getattr(L, "_adj_bound", False):
        return
    vp, i, d, dp, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
    ip = ctypes.POINTER(ctypes.c_int32)
    D = ctypes.POINTER(Desc)
    sig = {
        "swe_adj_create": (i, [vp, i64, i, ctypes.POINTER(vp)]),
        "swe_adj_create_levels": (i, [vp, i64, i64, i64, ctypes.POINTER(vp)]),
        "swe_adj_plan": (i, [i64, i64, i64, i64, ctypes.POINTER(ctypes.c_double)]),
        "swe_adj_destroy": (None, [vp]),
        "swe_adj_last_error": (ctypes.c_char_p, [vp]),
        "swe_adj_set_gauges": (i, [vp, i, ip, i, dp]),
        "swe_adj_forward": (i, [vp, dp, dp, dp]),
        "swe_adj_loss": (i, [vp, dp, dp, dp]),
        "swe_adj_backward": (i, [vp, dp, dp]),
        "swe_adj_get_series": (i, [vp, dp]),
        "swe_adj_get_dt": (i, [vp, dp]),
        "swe_adj_timing": (i, [vp, dp, dp]),
        "swe_adj_host_residual_rev": (i, [D, dp, dp, dp, dp, d, dp, dp, dp, dp, dp, dp]),
        "swe_adj_host_bc_rev": (i, [D, ip, dp, dp, dp, dp, dp, dp, dp, dp]),
        "swe_adj_host_bc_rev_gather": (i, [D, ip, dp, dp, dp, dp, dp, dp, dp, dp]),
        "swe_adj_host_bc_rev_onread": (i, [D, ip, dp, dp, dp, dp, dp, dp, dp, dp]),
        "swe_adj_host_epilogue_rev": (i, [D, i, dp, dp, dp, dp, dp, d, dp, dp, dp, dp, dp, dp, dp]),
        "swe_adj_host_hllc_rev": (None, [dp, d, d, d, d, d, d, d, d, d, d, d]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    L._adj_bound = True


def make_desc(nx, ny, rows, *, sphere=False, order=2, limiter="mc", g=9.80665, min_depth=1e-3, eps4=None,
              cfl=0.45, coriolis=0.0, manning=0.025, R=math.inf):
    """A swe_desc over host rows; returns (desc, kept arrays) -- keep both alive."""
    d = Desc()
    d.nx, d.ny, d.kind, d.order = nx, ny, int(sphere), order
    d.limiter = LIMITER[limiter] if isinstance(limiter, str) else limiter
    d.g, d.min_depth = g, min_depth
    d.eps4 = eps4 if eps4 is not None else js_pow4(max(min_depth, 1e-12))
    d.cfl, d.coriolis, d.manning, d.R = cfl, coriolis, manning, R
    keep = {r: np.ascontiguousarray(rows[r], dtype=np.float64) for r in ROWS}
    for k, r in enumerate(ROWS):
        d.rows[k] = _dp(keep[r])
    return d, keep


def hllc_rev(hL, huL, hvL, hR, huR, hvR, dry, g, lm, ln, lt):
    """Adjoint of hllc() (swe.mjs:974-1011) w.r.t. its six state arguments."""
    L = lib()
    _bind_adjoint(L)
    out = np.zeros(6)
    L.swe_adj_host_hllc_rev(_dp(out), hL, huL, hvL, hR, huR, hvR, dry, g, lm, ln, lt)
    return out


def residual_rev_host(desc, h, hu, hv, b, dt, lrh, lrhu, lrhv):
    """(d residual(dt) / d (h, hu, hv))^T (lrh, lrhu, lrhv) on the host, for a state with stored ghosts."""
    L = lib()
    _bind_adjoint(L)
    c = [np.ascontiguousarray(x, dtype=np.float64) for x in (h, hu, hv, b, lrh, lrhu, lrhv)]
    out = [np.zeros_like(c[0]) for _ in range(3)]
    rc = L.swe_adj_host_residual_rev(ctypes.byref(desc), *[_dp(x) for x in c[:4]], dt, *[_dp(x) for x in c[4:]],
                                     *[_dp(x) for x in out])
    if rc != 0:
        raise SweError(f"swe_adj_host_residual_rev: {rc}")
    return out


def bc_rev_host(desc, bc, lh, lhu, lhv, ext=None, state=None, gather=False, onread=False):
    """BC^T of applyBC() (swe.mjs:289-303) on copies; a flather side needs the
    far-field values ext (8) and the state (b, h, hu, hv) applyBC read.  gather:
    the form the device runs when no side is Flather and the south and north
    sides are not periodic; onread: the same folded on read, as the device's
    reverse epilogues apply it there."""
    L = lib()
    _bind_adjoint(L)
    kinds = (ctypes.c_int32 * 4)(*[BC[k] if isinstance(k, str) else k for k in bc])
    out = [np.array(x, dtype=np.float64, copy=True) for x in (lh, lhu, lhv)]
    e = None if ext is None else _ext(ext)
    st = [None] * 4 if state is None else [np.ascontiguousarray(x, dtype=np.float64) for x in state]
    ptr = lambda x: None if x is None else _dp(x)
    name = "swe_adj_host_bc_rev" + ("_onread" if onread else "_gather" if gather else "")
    fn = getattr(L, name)
    rc = fn(ctypes.byref(desc), kinds, ptr(e), ptr(st[1]), ptr(st[2]), ptr(st[3]), ptr(st[0]), *[_dp(x) for x in out])
    if rc != 0:
        raise SweError(f"{name}: {rc}")
    return out


def epilogue_rev_host(desc, stage, pre, nfield, dt, lout, sponge=None):
    """Adjoint of the pre-dryClean state at a stage's end (stage 2 also friction and
    the sponge with per-cell strengths `sponge`); returns (lpre_h, lpre_hu, lpre_hv, lN)."""
    L = lib()
    _bind_adjoint(L)
    p = [np.ascontiguousarray(x, dtype=np.float64) for x in pre]
    lo = [np.ascontiguousarray(x, dtype=np.float64) for x in lout]
    out = [np.zeros_like(p[0]) for _ in range(3)]
    lN = np.zeros_like(p[0])
    nf = None if nfield is None else np.ascontiguousarray(nfield, dtype=np.float64)
    sg = None if sponge is None else np.ascontiguousarray(sponge, dtype=np.float64)
    rc = L.swe_adj_host_epilogue_rev(ctypes.byref(desc), stage, *[_dp(x) for x in p],
                                     None if nf is None else _dp(nf), None if sg is None else _dp(sg), dt,
                                     *[_dp(x) for x in lo],
                                     *[_dp(x) for x in out], _dp(lN))
    if rc != 0:
        raise SweError(f"swe_adj_host_epilogue_rev: {rc}")
    return out[0], out[1], out[2], lN


def adjoint_plan_bytes(nsteps, ncells, inner, outer):
    """Device bytes of an adjoint of nsteps steps over ncells = W*H cells with
    replayed segments of `inner` steps and checkpoints every `outer` steps."""
    L = lib()
    _bind_adjoint(L)
    out = ctypes.c_double()
    rc = L.swe_adj_plan(int(nsteps), int(ncells), int(inner), int(outer), ctypes.byref(out))
    if rc != 0:
        raise SweError("swe_adj_plan: bad arguments")
    return out.value


class Adjoint:
    """Gradients of a gauge misfit through `nsteps` self-timed steps of a Solver
    (swe_adj_* of the C ABI): J = sum_s sum_g (eta_g(t_s) - obs[s, g])^2, eta
    sampled every `every` steps; dJ/dh0 (initial depth with zero momentum) and
    dJ/dn (per-cell Manning n).  The dt sequence is that of the last forward()."""

    def __init__(self, solver, nsteps, ckpt=0, outer=0):
        """ckpt: steps per replayed segment (0: chosen by the library); outer: steps
        between the checkpoints kept for the whole run, a multiple of ckpt (0: one
        level, outer = ckpt)."""
        L = lib()
        _bind_adjoint(L)
        self.solver = solver
        self.nsteps = int(nsteps)
        h = ctypes.c_void_p()
        if outer:
            rc = L.swe_adj_create_levels(solver._h, self.nsteps, int(ckpt), int(outer), ctypes.byref(h))
        else:
            rc = L.swe_adj_create(solver._h, self.nsteps, int(ckpt), ctypes.byref(h))
        if rc != 0:
            raise SweError(f"swe_adj_create: {L.swe_last_error(solver._h).decode()}")
        self._h = h
        self.ng, self.every = 0, 1

    def _ck(self, rc, what):
        if rc != 0:
            raise SweError(f"{what}: {lib().swe_adj_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            lib().swe_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_gauges(self, cells, every, obs):
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        self.ng, self.every = len(cells), int(every)
        self._ck(lib().swe_adj_set_gauges(self._h, len(cells), cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          int(every), _dp(obs)), "swe_adj_set_gauges")

    def _args(self, h0, nfield):
        h0 = np.ascontiguousarray(h0, dtype=np.float64)
        nf = None if nfield is None else np.ascontiguousarray(nfield, dtype=np.float64)
        return h0, nf

    def forward(self, h0, nfield=None):
        h0, nf = self._args(h0, nfield)
        J = ctypes.c_double()
        self._ck(lib().swe_adj_forward(self._h, _dp(h0), None if nf is None else _dp(nf), ctypes.byref(J)),
                 "swe_adj_forward")
        return J.value

    def loss(self, h0, nfield=None):
        h0, nf = self._args(h0, nfield)
        J = ctypes.c_double()
        self._ck(lib().swe_adj_loss(self._h, _dp(h0), None if nf is None else _dp(nf), ctypes.byref(J)),
                 "swe_adj_loss")
        return J.value

    def backward(self):
        gh = np.zeros(self.solver.N)
        gn = np.zeros(self.solver.N)
        self._ck(lib().swe_adj_backward(self._h, _dp(gh), _dp(gn)), "swe_adj_backward")
        return gh, gn

    def series(self):
        out = np.zeros((self.nsteps // self.every) * self.ng)
        self._ck(lib().swe_adj_get_series(self._h, _dp(out)), "swe_adj_get_series")
        return out.reshape(-1, self.ng) if self.ng else out

    def dts(self):
        out = np.zeros(self.nsteps)
        self._ck(lib().swe_adj_get_dt(self._h, _dp(out)), "swe_adj_get_dt")
        return out

    def timing(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        self._ck(lib().swe_adj_timing(self._h, ctypes.byref(f), ctypes.byref(b)), "swe_adj_timing")
        return f.value, b.value
