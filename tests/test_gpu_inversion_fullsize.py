"""The discrete adjoint at the inversion workload's full size: 1024x1024 cells,
dx = 50 m (shelf, 32 seamounts, wet/dry east coast, 32 gauges sampled every 10
steps), 240 self-timed steps (src/swe.mjs:754-814 reversed).

* The first guess is half the hidden truth surface plus 5 cm of smooth seeded
  noise (workloads.inversion_first_guess).
* Its forward is the solver's: the gauge series equals the oracle's host loop
  (the reference's step restated on the CPU) over the same 240 steps, to 1e-13
  (bit for bit in practice), the oracle running on a host thread meanwhile.
* dJ/d eta0 and dJ/dn agree with central finite differences of the loss along
  16 seeded directions of the control space (1 cm surface humps of 1-5 km
  centred within 2 km of a gauge, and a 1e-3 Fourier mode of the Manning field,
  on the wet cells), at steps 1e-3 to 1e-5 of the direction: a direction agrees
  when two successive steps give differences within 1e-6 of each other
  (step-converged) and the smaller step's is within 1e-6 of the adjoint.  At
  least 14 of the 16 must agree.
  Why near the gauges: in 240 steps a wave travels ~3 km of the 51 km domain,
  so a hump placed anywhere mostly lies outside what the gauges can see, and
  its directional derivative is small (1e-5 against 1e-2 here).  The loss is
  not differentiable everywhere -- a perturbation moves some limiter or wet/dry
  switch across its kink among 10^6 cells x 480 stage evaluations, where the
  adjoint takes the subgradient of the branch the forward took -- and for such
  a weak direction those kinks, and the loss's rounding noise (about 5e-14 of
  itself between two perturbed forwards), set a relative error of 1e-6 to 1e-4
  at any step; the humps placed anywhere on the domain were measured that way.
"""
import threading

import numpy as np
import pytest

import oracle
import swe_amd
import workloads

pytestmark = pytest.mark.gpu

NSTEPS, EVERY, NDIR = 240, 10, 16
STEPS = (1e-3, 1e-4, 1e-5)


@pytest.mark.timeout(900)
def test_inversion_adjoint_at_full_size():
    b, ht, p, gauges = workloads.inversion()
    nx, ny, dx = p["nx"], p["ny"], p["dx"]
    W, H = nx + 4, ny + 4
    rows = swe_amd.geometry_cartesian(nx, ny, dx, coriolis=p["coriolis"])
    z = np.zeros_like(ht)
    h_eval = workloads.inversion_first_guess(b, ht, nx, ny)
    X, Y = np.meshgrid(np.arange(W), np.arange(H))
    n_truth = (0.025 * (1 + 0.3 * np.sin(X / 97.0) * np.cos(Y / 131.0))).ravel()
    n0 = np.full(W * H, 0.025)

    # the oracle's host loop over the same steps, on a thread (ctypes releases the GIL)
    res = {}

    def run_oracle():
        o = oracle.OracleSim(nx, ny, rows, manning=p["manning"], coriolis=p["coriolis"], cfl=p["cfl"], dx=dx, dy=dx,
                             use_manning_field=True)
        o.load_state(b, h_eval, z, z)
        o.set_bc(p["bc"])
        o.set_manning_field(n0)
        ser = []
        for _ in range(NSTEPS // EVERY):
            o.run(EVERY)
            ser.append((o.b + o.h)[gauges])
        res["series"] = np.array(ser)
        res["t"] = o.t

    th = threading.Thread(target=run_oracle)
    th.start()

    s = swe_amd.Solver(nx, ny, rows, manning=p["manning"], coriolis=p["coriolis"], cfl=p["cfl"], dx=dx, dy=dx)
    s.load_state(b, h_eval, z, z)
    s.set_bc(p["bc"])
    s.set_manning_field(n0)
    adj = swe_amd.Adjoint(s, NSTEPS)
    nsamp = NSTEPS // EVERY
    adj.set_gauges(gauges, EVERY, np.zeros((nsamp, len(gauges))))
    adj.forward(ht, n_truth)
    adj.set_gauges(gauges, EVERY, adj.series())        # observations of the truth run
    J = adj.forward(h_eval, n0)
    series = adj.series()
    gh, gn = adj.backward()
    print(f"J = {J:.6e}, |dJ/dh0|max = {np.abs(gh).max():.3e}, |dJ/dn|max = {np.abs(gn).max():.3e}", flush=True)
    assert np.abs(gn).max() > 1e-6, "no friction signal in dJ/dn"

    wet = h_eval > 1.0
    agree, rows_out = 0, []
    for q, (dh, dn) in enumerate(workloads.inversion_directions(NDIR, nx, ny, dx, wet, seed=46, near=gauges)):
        ad = float(np.dot(gh, dh) + np.dot(gn, dn))
        fd = [(adj.loss(h_eval + e * dh, n0 + e * dn) - adj.loss(h_eval - e * dh, n0 - e * dn)) / (2 * e)
              for e in STEPS]
        ok = False
        for k in range(len(STEPS) - 1):
            conv = abs(fd[k] - fd[k + 1]) <= 1e-6 * abs(fd[k + 1])
            if conv and abs(fd[k + 1] - ad) <= 1e-6 * abs(ad):
                ok = True
        agree += ok
        rel = [abs(f - ad) / abs(ad) for f in fd]
        rows_out.append(rel)
        print(f"direction {q:2d}: adjoint {ad:+.10e}  rel err at " +
              "  ".join(f"{e:.0e}: {r:.2e}" for e, r in zip(STEPS, rel)) + ("  agrees" if ok else "  MISSES"),
              flush=True)
    th.join()
    dev = series.reshape(nsamp, len(gauges))
    ref = res["series"]
    err = np.abs(dev - ref).max() / np.abs(ref).max()
    print(f"gauge series vs the oracle: max rel {err:.2e} ({np.count_nonzero(dev != ref)} of {dev.size} differ); "
          f"{agree} of {NDIR} directions agree", flush=True)
    assert err <= 1e-13
    assert agree >= 14
