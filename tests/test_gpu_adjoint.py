"""The discrete adjoint on the GPU (swe_adj_* through the C ABI).

* its forward is the plain solver's run: the gauge series equals the oracle's
  run of the same steps (swe.mjs:754-814) sampled at the same cells;
* dJ/dh0 and dJ/dn agree with central finite differences of the forward that
  replays the recorded dt sequence (swe_adj_loss), to 1e-6 relative along
  smooth directions, and cell by cell (1e-6 for the kink-free schemes, 1e-4
  with MC/minmod limiters);
* the gradient does not depend on the checkpoint interval;
* Flather boundaries (device harmonics, swe.mjs:1084-1116) and the makeSponge
  relaxation (:1174-1198) are differentiated: FD cases and the host composition.
The states are rough (every surface difference well away from zero), so no
limiter sits on a kink where the finite difference and the adjoint's
subgradient would legitimately differ."""
import numpy as np
import pytest

import oracle
import swe_amd

pytestmark = pytest.mark.gpu


def rough_basin(nx, ny, dx, seed, *, wetdry=False):
    rng = np.random.default_rng(seed)
    W, H = nx + 4, ny + 4
    x = (np.arange(W) - 1.5) * dx
    y = (np.arange(H) - 1.5) * dx
    X, Y = np.meshgrid(x, y)
    L, M = nx * dx, ny * dx
    b = -4.0 + 0.5 * np.sin(2 * np.pi * Y / M) + rng.normal(0, 0.05, X.shape)
    if wetdry:
        b = b + 12.0 * np.maximum(0.0, X / L - 0.5)
    eta = (0.25 * np.exp(-((X - 0.35 * L) ** 2 + (Y - 0.45 * M) ** 2) / (0.02 * L * L))
           + 0.05 * np.sin(2 * np.pi * X / L) * np.cos(2 * np.pi * Y / M) + rng.normal(0, 0.01, X.shape))
    h = np.maximum(0.0, eta - b)
    return b.ravel(), h.ravel()


# the Flather forcing of the forced cases: a 0.15 m, 40 s wave with its incoming
# velocity, and a 0.05 m, 25 s one on the east side
HARMONIC = {0: dict(amp=[0.15], omega=[2 * np.pi / 40.0], phase=[0.3], uamp=[0.05], uomega=[2 * np.pi / 40.0],
                    uphase=[0.3]),
            1: dict(amp=[0.05], omega=[2 * np.pi / 25.0], phase=[1.1], msl=0.01)}


def sponge_band(nx, ny):
    s = np.zeros((ny + 4, nx + 4))
    s[2:-2, -10:-2] = np.linspace(0.002, 0.02, 8)
    return s.ravel()


def make(nx=40, ny=32, dx=25.0, *, order=2, limiter="mc", bc=("reflect",) * 4, coriolis=0.0, nsteps=48,
         every=4, ckpt=0, outer=0, wetdry=False, seed=1, harmonic=False, sponge=False):
    rows = swe_amd.geometry_cartesian(nx, ny, dx, coriolis=coriolis)
    s = swe_amd.Solver(nx, ny, rows, order=order, limiter=limiter, manning=0.0, coriolis=coriolis, dx=dx, dy=dx)
    b, h0 = rough_basin(nx, ny, dx, seed, wetdry=wetdry)
    z = np.zeros_like(h0)
    s.load_state(b, h0, z, z)
    s.set_bc(bc)
    if harmonic:
        for side, f in HARMONIC.items():
            if bc[side] == "flather":
                s.set_harmonic(side, **f)
    if sponge:
        s.set_sponge(sponge_band(nx, ny), eta_ref=0.02)
    rng = np.random.default_rng(seed + 100)
    nf = rng.uniform(0.015, 0.035, h0.size)
    W = nx + 4
    gauges = np.array([(4 + 7 * q) * W + 3 + (11 * q) % nx for q in range(4)] + [(ny // 2 + 2) * W + nx // 2 + 2],
                      dtype=np.int32)
    adj = swe_amd.Adjoint(s, nsteps, ckpt, outer)
    adj.set_gauges(gauges, every, np.zeros((nsteps // every, len(gauges))))
    adj.forward(h0, nf)
    ser = adj.series()
    obs = 0.6 * ser + 0.01 * np.sin(np.arange(ser.size)).reshape(ser.shape)
    adj.set_gauges(gauges, every, obs)
    ctx = dict(rows=rows, b=b, h0=h0, nf=nf, gauges=gauges, obs=obs, nx=nx, ny=ny, dx=dx, order=order,
               limiter=limiter, bc=bc, coriolis=coriolis, nsteps=nsteps, every=every, harmonic=harmonic,
               sponge=sponge)
    return s, adj, ctx


# Finite-difference steps: the forward's rounding noise in J is ~1e-14 absolute
# (the surface is b + h with |b|, |h| ~ 4 m), so a smaller perturbation loses the
# difference in noise; a larger one lets MC/minmod limiters change branch.  A
# 1e-8 m surface perturbation sits between the two.  n enters only friction,
# whose derivative along the n direction is small here (dJ ~1e-8 to 1e-7): a
# step of 1e-3 of the direction (1e-6 of n ~ 0.025) can leave its difference
# within 1e-6 of the noise, a step of 1e-2 can move a minmod limiter across its
# kink; the direction is held to the better of the two.
EPS_H, EPS_N = 1e-6, (1e-3, 1e-2)


def fd_dir(adj, h0, nf, dh=None, dn=None, eps=EPS_H):
    hp = h0 + eps * dh if dh is not None else h0
    hm = h0 - eps * dh if dh is not None else h0
    npl = nf + eps * dn if dn is not None else nf
    nmi = nf - eps * dn if dn is not None else nf
    return (adj.loss(hp, npl) - adj.loss(hm, nmi)) / (2 * eps)


def wet_interior(c, margin=0.2):
    W, H = c["nx"] + 4, c["ny"] + 4
    m = np.zeros((H, W), bool)
    m[2:-2, 2:-2] = True
    return m.ravel() & (c["h0"] > margin)


CASES = [
    dict(id="order1", order=1),
    dict(id="central", order=2, limiter="central"),
    dict(id="mc", order=2, limiter="mc"),
    dict(id="minmod-outflow", order=2, limiter="minmod", bc=("outflow", "reflect", "outflow", "reflect")),
    dict(id="mc-periodic-coriolis", order=2, limiter="mc", bc=("periodic",) * 4, coriolis=1e-3),
    dict(id="mc-flather", order=2, limiter="mc", bc=("flather", "reflect", "periodic", "periodic"), harmonic=True),
    dict(id="central-flather-both", order=2, limiter="central", bc=("flather", "flather", "reflect", "reflect"),
         harmonic=True),
    dict(id="mc-sponge", order=2, limiter="mc", bc=("reflect",) * 4, sponge=True),
    dict(id="order1-flather-sponge", order=1, bc=("flather", "reflect", "reflect", "outflow"), harmonic=True,
         sponge=True),
]


@pytest.mark.parametrize("case", CASES, ids=[c["id"] for c in CASES])
def test_adjoint_gradient_matches_finite_differences(case):
    kw = {k: v for k, v in case.items() if k != "id"}
    s, adj, c = make(**kw)
    J = adj.forward(c["h0"], c["nf"])
    assert J > 0
    gh, gn = adj.backward()
    assert np.isfinite(gh).all() and np.isfinite(gn).all()
    wet = wet_interior(c)
    rng = np.random.default_rng(5)
    W = c["nx"] + 4
    # smooth direction in h0 over wet interior cells
    X, Y = np.meshgrid(np.arange(W), np.arange(c["ny"] + 4))
    dh = (np.sin(X / 5.0) * np.cos(Y / 4.0)).ravel() * wet * 0.01
    fd = fd_dir(adj, c["h0"], c["nf"], dh=dh)
    ad = float(np.dot(gh, dh))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    # direction in n: a uniform change over the wet cells, the direction of the
    # largest derivative here (random and smooth patterns partly cancel, which
    # leaves their differences nearer the noise); the cell-by-cell checks below
    # probe single cells
    dn = wet * 1e-3
    fdn = [fd_dir(adj, c["h0"], c["nf"], dn=dn, eps=e) for e in EPS_N]
    adn = float(np.dot(gn, dn))
    assert min(abs(adn - f) / abs(f) for f in fdn) <= 1e-6, (adn, fdn)
    # cell by cell where the gradient is largest.  A one-cell perturbation spreads
    # over the whole basin within the run, so with the MC/minmod limiters some of
    # the ~10^6 limiter evaluations it touches change branch between the two
    # finite-difference runs; those cases are held to 1e-4, the kink-free schemes
    # (first order, central slopes) to 1e-6
    tol = 1e-6 if kw.get("order", 2) == 1 or kw.get("limiter") == "central" else 1e-4
    for k in np.argsort(-np.abs(gh) * wet)[:3]:
        e = np.zeros_like(gh)
        e[k] = 1.0
        fdk = fd_dir(adj, c["h0"], c["nf"], dh=e, eps=EPS_H)
        assert abs(gh[k] - fdk) <= tol * max(abs(fdk), 1e-3 * np.abs(gh).max()), (k, gh[k], fdk)
    for k in np.argsort(-np.abs(gn) * wet)[:3]:
        e = np.zeros_like(gn)
        e[k] = 1.0
        fdk = fd_dir(adj, c["h0"], c["nf"], dn=e, eps=1e-4)
        assert abs(gn[k] - fdk) <= tol * max(abs(fdk), 1e-3 * np.abs(gn).max()), (k, gn[k], fdk)


def test_adjoint_forward_is_the_solver_run():
    s, adj, c = make(order=2, limiter="mc", nsteps=24, every=1)
    adj.forward(c["h0"], c["nf"])
    ser = adj.series()
    dts = adj.dts()
    o = oracle.OracleSim(c["nx"], c["ny"], c["rows"], order=2, limiter="mc", manning=0.0, dx=c["dx"], dy=c["dx"],
                         use_manning_field=True)
    z = np.zeros_like(c["h0"])
    o.load_state(c["b"], c["h0"], z, z)
    o.set_manning_field(c["nf"])
    o.set_bc(c["bc"])
    for n in range(c["nsteps"]):
        dt = o.step()
        # friction's Math.pow may differ from the oracle's by an ulp; 1e-13 relative bounds it
        assert dt == pytest.approx(dts[n], rel=1e-13, abs=0)
        eta = o.b[c["gauges"]] + o.h[c["gauges"]]
        np.testing.assert_allclose(ser[n], eta, rtol=1e-13, atol=1e-13)


def test_adjoint_gradient_is_independent_of_checkpointing():
    grads = []
    for ckpt in (1, 5, 48):
        s, adj, c = make(order=2, limiter="mc", nsteps=48, ckpt=ckpt)
        adj.forward(c["h0"], c["nf"])
        grads.append(adj.backward())
    for gh, gn in grads[1:]:
        np.testing.assert_allclose(gh, grads[0][0], rtol=1e-12, atol=1e-12 * np.abs(grads[0][0]).max())
        np.testing.assert_allclose(gn, grads[0][1], rtol=1e-12, atol=1e-12 * np.abs(grads[0][1]).max())


def test_adjoint_two_levels_equal_one_level():
    # outer checkpoints every 12 / 9 / 15 steps, inner segments recomputed from
    # them (48 steps: whole, ragged and single-segment outer segments)
    grads = []
    for inner, outer in ((4, 0), (4, 12), (3, 9), (5, 15)):
        s, adj, c = make(order=2, limiter="mc", nsteps=48, ckpt=inner, outer=outer, wetdry=True)
        adj.forward(c["h0"], c["nf"])
        grads.append(adj.backward())
    for gh, gn in grads[1:]:
        np.testing.assert_allclose(gh, grads[0][0], rtol=1e-12, atol=1e-12 * np.abs(grads[0][0]).max())
        np.testing.assert_allclose(gn, grads[0][1], rtol=1e-12, atol=1e-12 * np.abs(grads[0][1]).max())


def test_adjoint_wet_dry_front():
    s, adj, c = make(order=2, limiter="mc", wetdry=True, nsteps=40)
    adj.forward(c["h0"], c["nf"])
    gh, gn = adj.backward()
    assert np.isfinite(gh).all() and np.isfinite(gn).all()
    W = c["nx"] + 4
    assert (c["h0"] <= 1e-3).any()
    wet = wet_interior(c, margin=0.5)
    X, Y = np.meshgrid(np.arange(W), np.arange(c["ny"] + 4))
    dh = (np.cos(X / 6.0) * np.sin(Y / 5.0)).ravel() * wet * 0.01
    fd = fd_dir(adj, c["h0"], c["nf"], dh=dh)
    ad = float(np.dot(gh, dh))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def test_adjoint_refusals():
    rows = swe_amd.geometry_cartesian(12, 10, 10.0)
    s = swe_amd.Solver(12, 10, rows, manning=0.0, dx=10.0, dy=10.0)
    s.set_bc(("stored", "reflect", "reflect", "reflect"))
    with pytest.raises(swe_amd.SweError, match="boundar"):
        swe_amd.Adjoint(s, 10)
    s.set_bc(("reflect",) * 4)
    adj = swe_amd.Adjoint(s, 10)
    with pytest.raises(swe_amd.SweError, match="forward"):
        adj.backward()
    with pytest.raises(swe_amd.SweError, match="interior"):
        adj.set_gauges(np.array([0], np.int32), 1, np.zeros((10, 1)))


HOST_CASES = [
    dict(id="mc-wetdry", limiter="mc", bc=("reflect",) * 4, wetdry=True),
    dict(id="minmod-mixed", limiter="minmod", bc=("outflow", "reflect", "periodic", "periodic")),
    dict(id="mc-flather-sponge", limiter="mc", bc=("flather", "flather", "periodic", "periodic"), harmonic=True,
         sponge=True),
    # several tiles of the tiled reverse passes along both axes, the last ones partial
    dict(id="mc-wide-wetdry", limiter="mc", bc=("periodic", "periodic", "reflect", "outflow"), wetdry=True,
         nx=150, ny=45),
]


def host_ext(t, bc):
    """The Flather values of HARMONIC at t, evaluated on the host."""
    e = np.zeros(8)
    for side, f in HARMONIC.items():
        if bc[side] != "flather":
            continue
        e[2 * side] = f.get("msl", 0.0) + sum(a * np.cos(w * t - p) for a, w, p in zip(f["amp"], f["omega"], f["phase"]))
        e[2 * side + 1] = sum(a * np.cos(w * t - p) for a, w, p in zip(f.get("uamp", []), f.get("uomega", []),
                                                                          f.get("uphase", [])))
    return e


@pytest.mark.parametrize("case", HOST_CASES, ids=[c["id"] for c in HOST_CASES])
def test_adjoint_step_equals_host_composition(case):
    # Two steps with one gauge sample after the second (so that the first step's
    # friction sees a momentum adjoint): the device's reverse (whose face flux
    # terms and positivity factors come from the replayed stage sweeps) against
    # the library's host reverse passes, which recompute everything, fed with the
    # oracle's stage states.  Equal up to the order of the atomic adds of the
    # boundary transpose.
    from test_adjoint_cpu import step_rev_composed
    nx, ny, dx = case.get("nx", 40), case.get("ny", 32), 25.0
    s, adj, c = make(nx, ny, dx, limiter=case["limiter"], bc=case["bc"], nsteps=2, every=2,
                     wetdry=case.get("wetdry", False), harmonic=case.get("harmonic", False),
                     sponge=case.get("sponge", False))
    adj.forward(c["h0"], c["nf"])
    gh, gn = adj.backward()
    dts = adj.dts()
    kw = dict(order=2, limiter=case["limiter"], manning=0.0)
    o = oracle.OracleSim(nx, ny, c["rows"], dx=dx, dy=dx, use_manning_field=True, **kw)
    desc, keep = swe_amd.make_desc(nx, ny, c["rows"], **kw)
    o.set_manning_field(c["nf"])
    o.set_bc(case["bc"])
    N = c["h0"].size
    z = np.zeros(N)
    strength = sponge_band(nx, ny) if case.get("sponge") else None
    if strength is not None:
        o.set_sponge(strength)
    e0, e1 = host_ext(0.0, case["bc"]), host_ext(float(dts[0]), case["bc"])
    o.load_state(c["b"], c["h0"], z, z)
    o.step(float(dts[0]), e0)
    if strength is not None:
        o.sponge(float(dts[0]), 0.02)
    A1 = np.array([o.h.copy(), o.hu.copy(), o.hv.copy()])
    o.step(float(dts[1]), e1)
    if strength is not None:
        o.sponge(float(dts[1]), 0.02)
    lam = np.zeros((3, N))
    for g, ob in zip(c["gauges"], c["obs"][0]):
        lam[0][g] += 2.0 * (c["b"][g] + o.h[g] - ob)
    lam, lN2 = step_rev_composed(o, desc, case["bc"], c["b"], A1, float(dts[1]), c["nf"], lam, ext=e1,
                                 sponge=strength)
    lA, lN1 = step_rev_composed(o, desc, case["bc"], c["b"], np.array([c["h0"], z, z]), float(dts[0]), c["nf"], lam,
                                ext=e0, sponge=strength)
    lN = lN2 + lN1
    assert np.abs(lA[0]).max() > 0 and np.abs(lN).max() > 0
    assert np.abs(gh - lA[0]).max() <= 1e-12 * np.abs(lA[0]).max()
    assert np.abs(gn - lN).max() <= 1e-12 * np.abs(lN).max()


@pytest.mark.parametrize("bc", [("reflect",) * 4, ("periodic", "periodic", "outflow", "reflect")])
def test_gradient_is_reproducible(bc):
    """With linear boundary kinds (no Flather side, south and north not periodic)
    every reverse pass writes each value once and BC^T gathers its ghosts in a
    fixed order, so two backward sweeps of one forward give the same gradient
    to the bit."""
    s, adj, c = make(bc=bc, wetdry=True, seed=5)
    adj.forward(c["h0"], c["nf"])
    g1 = [x.copy() for x in adj.backward()]
    adj.forward(c["h0"], c["nf"])
    g2 = adj.backward()
    for a, b in zip(g1, g2):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
