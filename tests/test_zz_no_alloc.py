"""swe_diagnostics_noalloc (the diagnostics entry point that allocates nothing,
its scratch in a static device array) against swe_diagnostics on the same
fields, through the C ABI: equal bit for bit, on two calls in a row, the second
on a smaller grid (the array keeps the first call's contents), and a third on
the sphere; a grid beyond the array's rows is refused."""
import ctypes

import numpy as np
import pytest

import swe_amd
from gpu_util import pair

pytestmark = pytest.mark.gpu


def _bind():
    L = swe_amd.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.swe_diagnostics_noalloc.restype = ctypes.c_int
    L.swe_diagnostics_noalloc.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32] + [ctypes.c_void_p] * 5 + \
        [ctypes.c_double] * 6 + [dp]
    return L


def dev_ptr(s, f):
    p = ctypes.POINTER(ctypes.c_double)()
    assert swe_amd.lib().swe_device_ptr(s._h, f, ctypes.byref(p)) == 0
    return ctypes.cast(p, ctypes.c_void_p).value


def noalloc(s, kind, area_ptr, dx, dy, ref):
    L = _bind()
    out = np.zeros(5)
    s.sync()
    rc = L.swe_diagnostics_noalloc(s.nx, s.ny, kind, *(dev_ptr(s, f) for f in (swe_amd.F_B, swe_amd.F_H,
                                                                               swe_amd.F_HU, swe_amd.F_HV)),
                                   area_ptr, dx, dy, s.min_depth, swe_amd.js_pow4(max(s.min_depth, 1e-12)), 9.80665,
                                   ref, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, L.swe_last_error(None)
    return out


def reference(s, ref):
    d = s.diagnostics(ref)
    return np.array([d["volume"], d["energy"], d["maxSurfaceDeviation"], d["maxSpeed"], float(d["finite"])])


def test_noalloc_diagnostics_equal_the_allocating_entry_point():
    print("[gpu] noalloc diagnostics: 1500x1100 then 130x90 Cartesian", flush=True)
    big, _ = pair(1500, 1100, 41, manning=0.025)
    for _ in range(3):
        big.step(None)
    want = reference(big, 0.05)
    got = noalloc(big, 0, None, 5.0, 5.0, 0.05)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
    small, _ = pair(130, 90, 42, manning=0.02)
    small.step(None)
    want = reference(small, -0.1)
    got = noalloc(small, 0, None, 5.0, 5.0, -0.1)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
    big.close()
    small.close()


def device_copy(a):
    """A device copy of a host array, through the HIP runtime the library runs on
    (test scaffolding: the caller's own device buffer)."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0   # host to device
    return hip, p


def test_noalloc_diagnostics_on_the_sphere_and_its_limit():
    print("[gpu] noalloc diagnostics: 180x90 sphere", flush=True)
    nx, ny = 180, 90
    rows = swe_amd.geometry_spherical(nx, ny, omega=7.292115e-5)
    s = swe_amd.Solver(nx, ny, rows, sphere=True, R=swe_amd.R_EARTH, manning=0.025)
    W, H = nx + 4, ny + 4
    lat, lon = swe_amd.sphere_lonlat(nx, ny)
    LA, LO = np.meshgrid(lat, lon, indexing="ij")
    b = (-4000 + 2500 * np.sin(3 * LO) * np.cos(2 * LA)).ravel()
    h = np.maximum(0.0, 1.5 * np.exp(-((LA - 0.3) / 0.2) ** 2).ravel() - b)
    z = np.zeros(W * H)
    s.load_state(b, h, z, z)
    for _ in range(4):
        s.step(None)
    hip, area = device_copy(np.ascontiguousarray(rows["area"], dtype=np.float64))
    want = reference(s, 0.0)
    got = noalloc(s, 1, area.value, 0.0, 0.0, 0.0)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
    assert hip.hipFree(area) == 0
    # a grid beyond the array's rows is refused before any kernel runs
    L = _bind()
    out = np.full(5, 7.0)
    rc = L.swe_diagnostics_noalloc(10, 16385, 0, dev_ptr(s, 0), dev_ptr(s, 1), dev_ptr(s, 2), dev_ptr(s, 3), None,
                                   1.0, 1.0, 1e-3, 1e-12, 9.80665, 0.0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == -1 and b"exceeds" in L.swe_last_error(None) and np.all(out == 7.0)
    s.close()
