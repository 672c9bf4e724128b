"""Source inversion with the adjoint, the use its gradient exists for (the
inversion workload: shelf, seamounts, east coast, gauges every 10 steps, here
at 256x256 with dx = 200 m so that waves reach the gauges in 400 steps).

The adjoint's forward/backward pair drives a standard quasi-Newton optimiser
(scipy's L-BFGS-B) over the initial surface of the wet cells, from a near-flat
sea: every evaluation is one recorded forward (the loss) and one backward
(dJ/d eta0).  The start is a flat sea plus 1 cm of noise (exactly flat, every
limiter argument sits on its kink).  The gauge misfit must fall by an order of
magnitude within 20 iterations and the estimate must move toward the hidden
truth surface."""
import numpy as np
import pytest
from scipy.optimize import minimize

import swe_amd
import workloads

pytestmark = pytest.mark.gpu


def test_lbfgs_on_the_adjoint_recovers_the_source():
    b, ht, p, gauges = workloads.inversion(nx=256, ny=256, dx=200.0, ngauges=32)
    nx, ny, dx = p["nx"], p["ny"], p["dx"]
    W, H = nx + 4, ny + 4
    rows = swe_amd.geometry_cartesian(nx, ny, dx, coriolis=p["coriolis"])
    s = swe_amd.Solver(nx, ny, rows, manning=p["manning"], coriolis=p["coriolis"], cfl=p["cfl"], dx=dx, dy=dx)
    z = np.zeros_like(ht)
    s.load_state(b, ht, z, z)
    s.set_bc(p["bc"])
    nsteps, every = 400, 10
    n0 = np.full(W * H, 0.025)
    adj = swe_amd.Adjoint(s, nsteps)
    adj.set_gauges(gauges, every, np.zeros((nsteps // every, len(gauges))))
    adj.forward(ht, n0)
    adj.set_gauges(gauges, every, adj.series())          # observations from the truth run
    inner = np.zeros((H, W), bool)
    inner[2:-2, 2:-2] = True
    wet = inner.ravel() & (b < -1.0)
    eta_truth = (b + ht)[wet]

    def h_of(e):
        h = np.maximum(0.0, -b)
        h[wet] = np.maximum(0.0, e - b[wet])
        return h

    calls = []

    def f(e):
        J = adj.forward(h_of(e), n0)
        gh, _ = adj.backward()
        calls.append(J)
        return J, gh[wet]                                  # dJ/d eta0 on the wet cells

    # a flat sea puts every limiter argument exactly on its kink (all surface
    # differences zero), where the subgradient is no descent direction: start
    # 1 cm of seeded noise away from it
    e0 = 0.01 * workloads.SplitMix64(47).normals(int(wet.sum()))
    J0 = f(e0)[0]
    r = minimize(f, e0, jac=True, method="L-BFGS-B", options=dict(maxiter=20))
    err0 = np.linalg.norm(e0 - eta_truth)
    err = np.linalg.norm(r.x - eta_truth)
    print(f"  {len(calls)} forward/backward pairs: misfit {J0:.4e} -> {r.fun:.4e}; "
          f"|eta - truth| {err0:.4e} -> {err:.4e}")
    assert r.fun < 0.1 * J0
    assert err < err0


def test_lbfgs_calibrates_manning():
    # friction calibration: a hump sloshing in a 3 m deep closed basin, gauges
    # every 10 steps over 600 steps; the hidden Manning field varies in space and
    # L-BFGS-B (n kept within [0.005, 0.08]) starts from a uniform 0.03
    nx, ny, dx = 64, 64, 20.0
    W, H = nx + 4, ny + 4
    rows = swe_amd.geometry_cartesian(nx, ny, dx)
    X, Y = np.meshgrid((np.arange(W) - 1.5) * dx, (np.arange(H) - 1.5) * dx)
    b = (-3.0 + 0.3 * np.sin(X / 150.0) * np.cos(Y / 170.0)).ravel()
    eta = (0.4 * np.exp(-((X - 400) ** 2 + (Y - 600) ** 2) / 150.0 ** 2)).ravel()
    h0 = np.maximum(0.0, eta - b)
    s = swe_amd.Solver(nx, ny, rows, manning=0.03, dx=dx, dy=dx)
    z = np.zeros_like(h0)
    s.load_state(b, h0, z, z)
    s.set_bc(("reflect",) * 4)
    nsteps, every = 600, 10
    gauges = np.array([(6 + 16 * (q // 4)) * W + 6 + 16 * (q % 4) for q in range(16)], dtype=np.int32)
    n_truth = (0.03 + 0.015 * np.sin(X / 200.0) * np.cos(Y / 260.0)).ravel()
    adj = swe_amd.Adjoint(s, nsteps)
    adj.set_gauges(gauges, every, np.zeros((nsteps // every, len(gauges))))
    adj.forward(h0, n_truth)
    adj.set_gauges(gauges, every, adj.series())
    inner = np.zeros((H, W), bool)
    inner[2:-2, 2:-2] = True
    inner = inner.ravel()
    calls = []

    def f(n_in):
        n = np.full(W * H, 0.03)
        n[inner] = n_in
        J = adj.forward(h0, n)
        _, gn = adj.backward()
        calls.append(J)
        return J, gn[inner]

    x0 = np.full(int(inner.sum()), 0.03)
    J0 = f(x0)[0]
    r = minimize(f, x0, jac=True, method="L-BFGS-B", bounds=[(0.005, 0.08)] * x0.size, options=dict(maxiter=20))
    err0 = np.linalg.norm(x0 - n_truth[inner])
    err = np.linalg.norm(r.x - n_truth[inner])
    print(f"  {len(calls)} forward/backward pairs: misfit {J0:.4e} -> {r.fun:.4e}; |n - truth| {err0:.4e} -> {err:.4e}")
    assert r.fun < 0.1 * J0
    assert err < err0
