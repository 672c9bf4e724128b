"""ctypes binding of the CPU oracle (oracle/swe_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing measured.
The oracle restates ShallowWater of src/swe.mjs (the reference solver); see the
header of swe_oracle.c for what pins each routine.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

B, H, HU, HV, MANNING, RH, RHU, RHV, THETA, SPONGE = range(10)
H0, HU0, HV0, U, V = 16, 17, 18, 19, 20
FXM, FXN, FXT, FYM, FYN, FYT, PXL, PXR, PYL, PYR = range(21, 31)
SE, SB, SU, SV = 40, 41, 42, 43
ROWS = ["dxRow", "dyRowN", "dyRowS", "dyCFL", "bedPhi", "tanPhi", "fRow", "geoCoef", "area"]
BC = {"reflect": 0, "outflow": 1, "periodic": 2, "flather": 3}
LIMITER = {"minmod": 0, "mc": 1, "central": 2}


class Params(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("sphere", ctypes.c_int32),
                ("order", ctypes.c_int32), ("limiter", ctypes.c_int32),
                ("use_manning_field", ctypes.c_int32),
                ("g", ctypes.c_double), ("min_depth", ctypes.c_double), ("eps4", ctypes.c_double),
                ("cfl", ctypes.c_double), ("coriolis", ctypes.c_double), ("manning", ctypes.c_double),
                ("R", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double)]


def build():
    """Compile liboracle.so in place (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        d, i, vp, dp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(dp)]
        L.orc_destroy.argtypes = [vp]
        L.orc_field.restype = dp
        L.orc_field.argtypes = [vp, i]
        L.orc_row.restype = dp
        L.orc_row.argtypes = [vp, i]
        for n in ("orc_W", "orc_H"):
            getattr(L, n).restype = i
            getattr(L, n).argtypes = [vp]
        L.orc_set_bc.argtypes = [vp, i, i]
        L.orc_set_time.argtypes = [vp, d, ctypes.c_int64]
        L.orc_time.restype = d
        L.orc_time.argtypes = [vp]
        L.orc_mass_floored.restype = d
        L.orc_mass_floored.argtypes = [vp]
        L.orc_apply_bc.argtypes = [vp, dp]
        L.orc_residual.argtypes = [vp, d]
        for n in ("orc_max_dt", "orc_dry_clean", "orc_volume", "orc_max_speed", "orc_energy"):
            getattr(L, n).restype = d
            getattr(L, n).argtypes = [vp]
        L.orc_max_surface_deviation.restype = d
        L.orc_max_surface_deviation.argtypes = [vp, d]
        L.orc_finite.restype = i
        L.orc_finite.argtypes = [vp]
        L.orc_apply_friction.argtypes = [vp, d]
        L.orc_step.restype = d
        L.orc_step.argtypes = [vp, d, dp]
        L.orc_sponge.argtypes = [vp, d, d]
        L.orc_run.restype = ctypes.c_int64
        L.orc_run.argtypes = [vp, ctypes.c_int64, dp]
        L.orc_hllc.argtypes = [dp] + [d] * 8
        L.orc_vel.restype = d
        L.orc_vel.argtypes = [d, d, d]
        L.orc_pow.restype = d
        L.orc_pow.argtypes = [d, d]
        L.orc_hypot.restype = d
        L.orc_hypot.argtypes = [d, d]
        L.orc_limit.restype = d
        L.orc_limit.argtypes = [i, d, d]
        L.orc_fnv1a.restype = ctypes.c_uint64
        L.orc_fnv1a.argtypes = [vp, ctypes.c_size_t, ctypes.c_uint64]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ext_array(ext):
    e = np.zeros(8, dtype=np.float64)
    if ext is not None:
        e[:] = ext
    return e


class OracleSim:
    """One oracle solver instance; fields are numpy views of the C arrays."""

    def __init__(self, nx, ny, rows, *, sphere=False, order=2, limiter="mc", g=9.80665,
                 min_depth=1e-3, eps4=None, cfl=0.45, coriolis=0.0, manning=0.025,
                 R=math.inf, dx=1.0, dy=None, use_manning_field=False):
        L = lib()
        if eps4 is None:
            eps4 = L.orc_pow(max(min_depth, 1e-12), 4.0)
        self.p = Params(nx, ny, int(sphere), order, LIMITER[limiter] if isinstance(limiter, str) else limiter,
                        int(use_manning_field), g, min_depth, eps4, cfl, coriolis, manning, R, dx,
                        dx if dy is None else dy)
        self.nx, self.ny = nx, ny
        self.W, self.H = nx + 4, ny + 4
        self.N = self.W * self.H
        self._rows = [np.ascontiguousarray(rows[r], dtype=np.float64) for r in ROWS]
        arr = (ctypes.POINTER(ctypes.c_double) * len(ROWS))(*[_dp(r) for r in self._rows])
        self.h_ = L.orc_create(ctypes.byref(self.p), arr)
        if not self.h_:
            raise MemoryError("orc_create failed")

    def __del__(self):
        if getattr(self, "h_", None):
            lib().orc_destroy(self.h_)
            self.h_ = None

    def field(self, which):
        n = 2 * self.N if which >= SE else self.N
        ptr = lib().orc_field(self.h_, which)
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def row(self, which):
        return np.ctypeslib.as_array(lib().orc_row(self.h_, which), shape=(self.H,))

    b = property(lambda s: s.field(B))
    h = property(lambda s: s.field(H))
    hu = property(lambda s: s.field(HU))
    hv = property(lambda s: s.field(HV))

    def load_state(self, b, h, hu, hv):
        self.b[:] = b
        self.h[:] = h
        self.hu[:] = hu
        self.hv[:] = hv

    def state(self):
        return self.b.copy(), self.h.copy(), self.hu.copy(), self.hv.copy()

    def set_manning_field(self, n):
        self.field(MANNING)[:] = n
        self.p.use_manning_field = 1
        # the C side holds its own copy of the parameters: rebuild it
        L = lib()
        saved = [self.field(f).copy() for f in (B, H, HU, HV, MANNING, SPONGE)]
        t = self.t
        L.orc_destroy(self.h_)
        arr = (ctypes.POINTER(ctypes.c_double) * len(ROWS))(*[_dp(r) for r in self._rows])
        self.h_ = L.orc_create(ctypes.byref(self.p), arr)
        for f, a in zip((B, H, HU, HV, MANNING, SPONGE), saved):
            self.field(f)[:] = a
        self.set_time(t)
        if hasattr(self, "_bc"):
            self.set_bc(self._bc)

    def set_sponge(self, strength):
        self.field(SPONGE)[:] = strength

    def set_bc(self, kinds):
        self._bc = list(kinds)
        for side, k in enumerate(kinds):
            lib().orc_set_bc(self.h_, side, BC[k] if isinstance(k, str) else k)

    def apply_bc(self, ext=None):
        e = ext_array(ext)
        lib().orc_apply_bc(self.h_, _dp(e))

    def residual(self, dt):
        lib().orc_residual(self.h_, dt)

    def max_dt(self):
        return lib().orc_max_dt(self.h_)

    def step(self, dt=None, ext=None):
        e = ext_array(ext)
        return lib().orc_step(self.h_, math.nan if dt is None else dt, _dp(e))

    def run(self, nsteps, ext=None):
        e = ext_array(ext)
        return lib().orc_run(self.h_, nsteps, _dp(e))

    def sponge(self, dt, ref):
        lib().orc_sponge(self.h_, dt, ref)

    def apply_friction(self, dt):
        lib().orc_apply_friction(self.h_, dt)

    def dry_clean(self):
        return lib().orc_dry_clean(self.h_)

    @property
    def t(self):
        return lib().orc_time(self.h_)

    def set_time(self, t, steps=0):
        lib().orc_set_time(self.h_, t, steps)

    @property
    def mass_floored(self):
        return lib().orc_mass_floored(self.h_)

    def volume(self):
        return lib().orc_volume(self.h_)

    def max_speed(self):
        return lib().orc_max_speed(self.h_)

    def energy(self):
        return lib().orc_energy(self.h_)

    def max_surface_deviation(self, ref):
        return lib().orc_max_surface_deviation(self.h_, ref)

    def finite(self):
        return bool(lib().orc_finite(self.h_))

    def fingerprint(self):
        L = lib()
        hsh = 0
        for a in (self.h, self.hu, self.hv):
            hsh = L.orc_fnv1a(a.ctypes.data, a.nbytes, hsh)
        return hsh


def hllc(hL, huL, hvL, hR, huR, hvR, dry=1e-4, g=9.80665):
    out = np.zeros(3)
    lib().orc_hllc(_dp(out), hL, huL, hvL, hR, huR, hvR, dry, g)
    return out


def vel(hq, h, eps4):
    return lib().orc_vel(hq, h, eps4)


def js_pow(x, y):
    return lib().orc_pow(x, y)


def js_hypot(a, b):
    return lib().orc_hypot(a, b)


def limit(lim, a, b):
    return lib().orc_limit(LIMITER[lim] if isinstance(lim, str) else lim, a, b)
