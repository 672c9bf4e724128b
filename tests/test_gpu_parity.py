"""Parity of the HIP step (through the C ABI) with the reference and the oracle.

Bar: bit-exact.  Every golden scenario is replayed on the GPU and must match the
reference solver's fingerprints exactly wherever the oracle does, and must match
the oracle itself bit for bit at every step in all scenarios (friction included:
both use the same pow).  Seeded random basins of several shapes -- crossing the
58-column strip and row-chunk boundaries of the fused kernel -- are compared
with the oracle bit for bit over 25 steps for every boundary kind, limiter,
order, friction mode and the sponge.
"""
import math
import sys

import numpy as np
import pytest

import goldens
import oracle
import swe_amd
from gpu_util import bits, first_diff, pair

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_factory(nx, ny, rows, **kw):
    kw.pop("use_manning_field", None)
    return swe_amd.Solver(nx, ny, rows, **kw)


@pytest.mark.parametrize("name", goldens.STEP_SCENARIOS)
def test_gpu_replays_reference(name):
    say(f"[gpu] replay {name}")
    fx = goldens.load(name)
    sim = goldens.make_sim(gpu_factory, fx)
    fr = goldens.uses_friction(fx)
    problems = goldens.replay(sim, fx, exact=not fr, rtol=1e-13)
    assert not problems, problems


@pytest.mark.parametrize("name", goldens.STEP_SCENARIOS)
def test_gpu_matches_oracle_every_step(name):
    say(f"[gpu] vs oracle {name}")
    fx = goldens.load(name)
    g = goldens.make_sim(gpu_factory, fx)
    o = goldens.make_sim(oracle.OracleSim, fx)
    sponge = fx.get("sponge")
    W = fx["params"]["nx"] + 4
    for n, st in enumerate(fx["steps"], start=1):
        ext = [goldens.hexf(x) for x in st["ext"]]
        dt_arg = None if st["dtArg"] is None else goldens.hexf(st["dtArg"])
        dg = g.step(dt_arg, ext)
        do = o.step(dt_arg, ext)
        if sponge and do > 0:
            o.sponge(do, goldens.hexf(sponge["etaRef"]))
        if not st["finite"]:
            assert not g.finite()
            break
        assert dg == do, (n, dg, do)
        _, h, hu, hv = g.state()
        for nm, a, b in (("h", h, o.h), ("hu", hu, o.hu), ("hv", hv, o.hv)):
            assert first_diff(a, b, W) is None, (n, nm, first_diff(a, b, W))
        assert g.mass_floored == o.mass_floored or abs(g.mass_floored - o.mass_floored) <= 1e-15 * abs(o.mass_floored)


@pytest.mark.parametrize("name", ["probe_positivity", "probe_sphere"])
def test_gpu_residual_probes(name):
    say(f"[gpu] probe {name}")
    fx = goldens.load(name)
    for pr in fx["probes"]:
        sim = goldens.make_sim(gpu_factory, {**fx, "manningField": None, "sponge": None})
        assert sim.max_dt() == goldens.hexf(fx["maxDt"])
        sim.apply_bc()
        b, h, hu, hv = sim.state()
        for key, a in (("b", b), ("h", h), ("hu", hu), ("hv", hv)):
            assert np.array_equal(bits(a), bits(goldens.arr(pr["after_bc"][key]))), key
        sim.residual(goldens.hexf(pr["dt"]))
        for f, key in ((swe_amd.F_RH, "rh"), (swe_amd.F_RHU, "rhu"), (swe_amd.F_RHV, "rhv"),
                       (swe_amd.F_THETA, "theta")):
            got, want = sim.get_field(f), goldens.arr(pr[key])
            W = fx["params"]["nx"] + 4
            assert first_diff(got, want, W) is None, (pr["mult"], key, first_diff(got, want, W))


def test_gpu_stage2_theta_and_rh_after_step():
    say("[gpu] theta/rh after step")
    for name in ("lake_island", "dambreak_dry", "sphere_pole"):
        fx = goldens.load(name)
        sim = goldens.make_sim(gpu_factory, fx)
        st = fx["steps"][0]
        sim.step(None, [goldens.hexf(x) for x in st["ext"]], want_theta=True, want_rh=True)
        cp = fx["checkpoints"]["1"]
        W = fx["params"]["nx"] + 4
        for f, key in ((swe_amd.F_THETA, "theta"), (swe_amd.F_RH, "rh"), (swe_amd.F_RHU, "rhu"),
                       (swe_amd.F_RHV, "rhv"), (swe_amd.F_B, "b")):
            got = sim.get_field(f)
            assert first_diff(got, goldens.arr(cp[key]), W) is None, (name, key)


CASES = [
    dict(nx=100, ny=70, bc=("reflect",) * 4),
    dict(nx=131, ny=61, bc=("outflow", "reflect", "periodic", "periodic")),
    dict(nx=57, ny=150, bc=("periodic", "periodic", "reflect", "outflow"), limiter="minmod"),
    dict(nx=120, ny=40, bc=("flather", "flather", "reflect", "flather"), manning=0.03),
    dict(nx=117, ny=33, bc=("reflect",) * 4, order=1, coriolis=2e-3),
    dict(nx=64, ny=64, bc=("periodic",) * 4, limiter="central", wetdry=False),
    dict(nx=90, ny=45, bc=("reflect", "outflow", "reflect", "reflect"), manning_field=True, sponge=True),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_gpu_random_basins_bit_exact(case):
    kw = dict(CASES[case])
    nx, ny = kw.pop("nx"), kw.pop("ny")
    say(f"[gpu] random basin {case} {nx}x{ny} {kw}")
    g, o = pair(nx, ny, 1000 + case, **kw)
    W = nx + 4
    ext = [0.3, 0.02, -0.1, 0.0, 0.0, 0.0, 0.05, -0.01]
    for chunk in (None,):
        pass
    for n in range(25):
        dg = g.step(None, ext)
        do = o.step(None, ext)
        if kw.get("sponge"):
            o.sponge(do, 0.0)
        assert dg == do, (n, dg, do)
        _, h, hu, hv = g.state()
        for nm, a, b in (("h", h, o.h), ("hu", hu, o.hu), ("hv", hv, o.hv)):
            assert first_diff(a, b, W) is None, (n, nm, first_diff(a, b, W))


def test_gpu_uneven_cartesian_rows_bit_exact():
    """A Cartesian grid whose rows differ (x spacing, both y spacings, the CFL and
    bed divisors and f varying smoothly by row, as a stretched grid's would) runs
    the stage kernels that read every row's metric but leave out the spherical
    terms; over 25 steps it equals the oracle bit for bit."""
    from gpu_util import random_basin
    nx, ny, dx = 97, 66, 5.0
    say(f"[gpu] uneven Cartesian rows {nx}x{ny}")
    rows = dict(swe_amd.geometry_cartesian(nx, ny, dx, coriolis=1e-3))
    j = np.arange(ny + 4)
    for key, amp in (("dxRow", 0.03), ("dyRowN", 0.02), ("dyRowS", 0.02), ("dyCFL", 0.02), ("bedPhi", 0.02),
                     ("fRow", 0.1)):
        rows[key] = np.array(rows[key], dtype=np.float64) * (1.0 + amp * np.sin(0.3 * j + len(key)))
    kw = dict(order=2, limiter="mc", manning=0.02, coriolis=1e-3, dx=dx, dy=dx)
    g = swe_amd.Solver(nx, ny, rows, **kw)
    o = oracle.OracleSim(nx, ny, rows, **kw)
    st = random_basin(nx, ny, 77, dx=dx)
    bc = ("flather", "reflect", "periodic", "periodic")
    for x in (g, o):
        x.load_state(*st)
        x.set_bc(bc)
    W = nx + 4
    ext = [0.2, 0.01, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    for n in range(25):
        dg = g.step(None, ext)
        do = o.step(None, ext)
        assert dg == do, (n, dg, do)
        _, h, hu, hv = g.state()
        for nm, a, b in (("h", h, o.h), ("hu", hu, o.hu), ("hv", hv, o.hv)):
            assert first_diff(a, b, W) is None, (n, nm, first_diff(a, b, W))


@pytest.mark.parametrize("rows", [8, 13, 40, 1000])
def test_gpu_chunk_rows_do_not_change_results(rows):
    say(f"[gpu] chunk rows {rows}")
    g1, _ = pair(150, 90, 7)
    g2, _ = pair(150, 90, 7)
    g2.set_chunk_rows(rows)
    for _ in range(10):
        g1.step(None)
        g2.step(None)
    for a, b in zip(g1.state(), g2.state()):
        assert np.array_equal(bits(a), bits(b))


def test_gpu_advance_and_graph_match_strict_steps():
    say("[gpu] advance / graph")
    g1, _ = pair(200, 120, 11, bc=("flather", "reflect", "reflect", "reflect"), manning=0.025)
    g2, _ = pair(200, 120, 11, bc=("flather", "reflect", "reflect", "reflect"), manning=0.025)
    g3, _ = pair(200, 120, 11, bc=("flather", "reflect", "reflect", "reflect"), manning=0.025)
    for g in (g2, g3):
        g.set_harmonic(0, amp=[0.0], omega=[0.0], phase=[0.0], msl=0.0)
    g3.use_graph(True)
    for _ in range(9):
        g1.step(None, [0.0] * 8)
    assert g2.advance(9) == 9
    assert g3.advance(9) == 9
    for a, b, c in zip(g1.state(), g2.state(), g3.state()):
        assert np.array_equal(bits(a), bits(b))
        assert np.array_equal(bits(a), bits(c))
    assert g1.t == g2.t == g3.t


def test_gpu_refused_step_keeps_state():
    say("[gpu] refused step")
    g, o = pair(40, 30, 3)
    before = g.state()
    assert g.step(0.0) == 0.0
    assert g.step(math.inf) == 0.0
    o.apply_bc()
    after = g.state()
    for a, b in zip(after[1:], (o.h, o.hu, o.hv)):
        assert np.array_equal(bits(a), bits(b))
    assert np.array_equal(bits(after[0]), bits(o.b))
    _ = before


def test_gpu_all_dry_max_dt_is_infinite():
    rows = swe_amd.geometry_cartesian(20, 10, 1.0)
    g = swe_amd.Solver(20, 10, rows, manning=0.0)
    assert g.max_dt() == math.inf
    assert g.step(None) == 0.0


def test_gpu_short_chunks_bit_identical():
    """Chunks of 1 to 6 rows (what the autotuner tries on small grids) give the
    8-row chunk's results bit for bit, on a wet/dry basin and on the sphere."""
    say("[gpu] short chunks")
    ref_state = None
    for rows in (8, 1, 2, 3, 5, 6):
        g, _ = pair(131, 45, 21, bc=("flather", "reflect", "periodic", "periodic"), manning=0.02)
        g.set_chunk_rows(rows)
        for _ in range(12):
            g.step(None, [0.1, 0.0] + [0.0] * 6)
        st = [bits(a) for a in g.state()] + [g.t]
        if ref_state is None:
            ref_state = st
        else:
            assert all(np.array_equal(a, b) for a, b in zip(st[:4], ref_state[:4])) and st[4] == ref_state[4], rows
    nx, ny = 72, 36
    rows_s = swe_amd.geometry_spherical(nx, ny)
    lat = np.repeat(swe_amd.sphere_lonlat(nx, ny)[0], nx + 4)
    b = -4000.0 + 800.0 * np.cos(3 * lat)
    h = np.maximum(0.0, 1.0 * np.sin(lat) ** 2 - b)
    z = np.zeros_like(h)
    out = []
    for rows in (8, 1, 3):
        s = swe_amd.Solver(nx, ny, rows_s, sphere=True, R=swe_amd.R_EARTH, manning=0.02)
        s.load_state(b, h, z, z)
        s.set_bc(("periodic", "periodic", "reflect", "reflect"))
        s.set_chunk_rows(rows)
        assert s.advance(20) == 20
        out.append([bits(a) for a in s.state()[1:]])
        s.close()
    for o in out[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(o, out[0]))
