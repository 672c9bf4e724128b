"""The CPU oracle against the reference solver's own outputs (golden fixtures).

Non-friction scenarios must agree bit for bit at every step: dt (so maxDt),
massFloored and the FNV-1a fingerprint of the raw bytes of h, hu, hv including
the ghost ring.  Scenarios through applyFriction are compared at 1e-13 relative
with an identical wet/dry mask, because the engine's Math.pow is restated only
to 1 ulp (orc_pow).
"""
import numpy as np
import pytest

import goldens
import oracle


@pytest.mark.parametrize("name", goldens.STEP_SCENARIOS)
def test_oracle_replays_reference(name):
    fx = goldens.load(name)
    sim = goldens.make_sim(oracle.OracleSim, fx)
    fr = goldens.uses_friction(fx)
    problems = goldens.replay(sim, fx, exact=not fr, rtol=1e-13)
    assert not problems, problems


@pytest.mark.parametrize("name", ["manning_scalar", "sphere_band_o1"])
def test_friction_scenarios_that_happen_to_be_exact(name):
    # these two never hit an argument where orc_pow and the engine's pow differ
    fx = goldens.load(name)
    sim = goldens.make_sim(oracle.OracleSim, fx)
    assert not goldens.replay(sim, fx, exact=True)


@pytest.mark.parametrize("name", ["probe_positivity", "probe_sphere"])
def test_residual_probes(name):
    """applyBC(); residual(dt) -- rh, rhu, rhv, theta and face mass fluxes, bit for bit."""
    fx = goldens.load(name)
    for pr in fx["probes"]:
        sim = goldens.make_sim(oracle.OracleSim, {**fx, "manningField": None, "sponge": None})
        assert sim.max_dt() == goldens.hexf(fx["maxDt"])
        sim.apply_bc()
        for f, key in ((oracle.B, "b"), (oracle.H, "h"), (oracle.HU, "hu"), (oracle.HV, "hv")):
            assert np.array_equal(sim.field(f), goldens.arr(pr["after_bc"][key])), key
        sim.residual(goldens.hexf(pr["dt"]))
        for f, key in ((oracle.RH, "rh"), (oracle.RHU, "rhu"), (oracle.RHV, "rhv"),
                       (oracle.THETA, "theta"), (oracle.FXM, "fxM"), (oracle.FYM, "fyM")):
            got, want = sim.field(f), goldens.arr(pr[key])
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (pr["mult"], key)


def test_positivity_probe_fires():
    # the probe at 5x and 20x CFL must actually exercise theta < 1
    fx = goldens.load("probe_positivity")
    fired = [float(np.min(goldens.arr(p["theta"]))) for p in fx["probes"]]
    assert min(fired) < 1.0


def test_checkpoint_arrays_first_step():
    """Full arrays after step 1, including rh/theta of stage 2 and ghost cells."""
    for name in ("lake_island", "flather_tide", "sphere_pole", "dambreak_dry"):
        fx = goldens.load(name)
        sim = goldens.make_sim(oracle.OracleSim, fx)
        st = fx["steps"][0]
        sim.step(None, [goldens.hexf(x) for x in st["ext"]])
        cp = fx["checkpoints"]["1"]
        for f, key in ((oracle.H, "h"), (oracle.HU, "hu"), (oracle.HV, "hv"), (oracle.B, "b"),
                       (oracle.RH, "rh"), (oracle.RHU, "rhu"), (oracle.RHV, "rhv"), (oracle.THETA, "theta")):
            assert np.array_equal(sim.field(f).view(np.uint64), goldens.arr(cp[key]).view(np.uint64)), (name, key)


def test_diagnostics_match_reference():
    for name in ("lake_island", "sphere_pole", "periodic_all"):
        fx = goldens.load(name)
        sim = goldens.make_sim(oracle.OracleSim, fx)
        v0 = sim.volume()
        assert v0 == goldens.hexf(fx["final"]["volume0"])
        goldens.replay(sim, fx)
        fin = fx["final"]
        assert sim.volume() == goldens.hexf(fin["volume"])
        assert sim.max_speed() == goldens.hexf(fin["maxSpeed"])
        assert sim.energy() == goldens.hexf(fin["energy"])
        assert sim.max_surface_deviation(0.0) == goldens.hexf(fin["maxSurfaceDeviation0"])
        assert sim.finite() == fin["finite"]


def test_planted_gravity_mistake_is_caught():
    """The golden comparison can fail: g one part in 1e9 off breaks the first step."""
    fx = goldens.load("wavy_coriolis")
    kw = goldens.params(fx)
    nx, ny = kw.pop("nx"), kw.pop("ny")
    kw["g"] = kw["g"] * (1 + 1e-9)
    sim = oracle.OracleSim(nx, ny, goldens.rows(fx), **kw)
    sim.set_bc(fx["bc"])
    init = {k: goldens.arr(v) for k, v in fx["init"].items()}
    sim.load_state(init["b"], init["h"], init["hu"], init["hv"])
    assert goldens.replay(sim, fx)


def test_planted_one_ulp_state_nudge_is_caught():
    fx = goldens.load("lake_island")
    sim = goldens.make_sim(oracle.OracleSim, fx)
    k = sim.W * 10 + 10
    sim.h[k] = np.nextafter(sim.h[k], np.inf)
    assert goldens.replay(sim, fx)


def test_oracle_replays_thacker256_two_periods():
    """The reference's whole 2088-step Thacker run (256x256, dt = min(maxDt, 2T - t)):
    every step's dt and fingerprint, bit for bit (about a minute of CPU)."""
    fx = goldens.load("thacker256")
    sim = goldens.make_sim(oracle.OracleSim, fx)
    for n, st in enumerate(fx["steps"], start=1):
        assert sim.step(goldens.hexf(st["dtArg"])) == goldens.hexf(st["dt"]), n
        assert goldens.fingerprint(sim) == int(st["hash"], 16), n
    assert sim.t == goldens.hexf(fx["final"]["t"])
