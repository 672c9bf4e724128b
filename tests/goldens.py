"""Loading and replaying the golden fixtures written by tests/golden/make_golden.mjs.

A fixture holds the reference solver's inputs (initial arrays, geometry rows,
parameters, boundary kinds, per-step boundary forcing) and its outputs
(per-step dt, massFloored, an FNV-1a-64 fingerprint of h,hu,hv, checkpoints).
`replay()` drives any solver object exposing the oracle's method names
(OracleSim, or the HIP host mirror) through the same steps and reports where
it first departs from the reference.
"""
import base64
import gzip
import json
import math
import os
import struct

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROWS = ["dxRow", "dyRowN", "dyRowS", "dyCFL", "bedPhi", "tanPhi", "fRow", "geoCoef", "area"]
STEP_SCENARIOS = ["lake_island", "dambreak_dry", "wavy_coriolis", "order1_island", "minmod_beach",
                  "central_wave", "periodic_all", "sponge_wave", "flather_tide", "sphere_pole",
                  "manning_scalar", "manning_field", "sphere_band_o1", "drying_flat", "blowup"]


def hexf(s):
    return struct.unpack(">d", bytes.fromhex(s))[0]


def arr(s):
    return np.frombuffer(base64.b64decode(s), dtype="<f8").copy()


def load(name):
    with gzip.open(os.path.join(GOLDEN_DIR, name + ".json.gz"), "rt") as f:
        return json.load(f)


def available():
    return sorted(n[:-8] for n in os.listdir(GOLDEN_DIR) if n.endswith(".json.gz"))


def params(fx):
    p = fx["params"]
    sph = p["kind"] == "sphere"
    return dict(
        nx=p["nx"], ny=p["ny"], sphere=sph, order=p["order"], limiter=p["limiter"],
        g=hexf(fx["G"]), min_depth=hexf(p["minDepth"]), eps4=hexf(p["eps4"]), cfl=hexf(p["cfl"]),
        coriolis=hexf(p["coriolis"]), manning=hexf(p["manning"]), R=hexf(p["R"]),
        dx=(hexf(p["dx"]) if p["dx"] is not None else math.nan), dy=hexf(p["dy"]),
        use_manning_field=fx.get("manningField") is not None)


def rows(fx):
    return {r: arr(fx["rows"][r]) for r in ROWS}


def uses_friction(fx):
    p = fx["params"]
    return fx.get("manningField") is not None or hexf(p["manning"]) != 0.0


def make_sim(factory, fx):
    """factory(nx, ny, rows, **params) -> solver; loads the initial state."""
    kw = params(fx)
    nx, ny = kw.pop("nx"), kw.pop("ny")
    sim = factory(nx, ny, rows(fx), **kw)
    sim.set_bc(fx["bc"])
    init = {k: arr(v) for k, v in fx["init"].items()}
    sim.load_state(init["b"], init["h"], init["hu"], init["hv"])
    if fx.get("manningField") is not None:
        sim.set_manning_field(arr(fx["manningField"]))
    if fx.get("sponge"):
        sim.set_sponge(arr(fx["sponge"]["strength_probe"]))
    return sim


def fingerprint(sim):
    """FNV-1a-64 of the raw bytes of h, hu, hv (the C routine of the oracle library)."""
    import oracle
    L = oracle.lib()
    hsh = 0
    for a in sim.state()[1:]:
        a = np.ascontiguousarray(a)
        hsh = L.orc_fnv1a(a.ctypes.data, a.nbytes, hsh)
    return hsh


def replay(sim, fx, *, exact=True, rtol=1e-12, on_step=None):
    """Step `sim` through fixture `fx`.  Returns a list of problems (empty = parity).

    exact=True: dt, massFloored and the FNV-1a fingerprint must match bit for bit.
    exact=False: dt and the checkpoint arrays must agree to rtol (relative to the
    field's max magnitude) and the wet/dry mask (h <= minDepth) must be identical.
    """
    problems = []
    md = hexf(fx["params"]["minDepth"])
    sponge = fx.get("sponge")
    for n, st in enumerate(fx["steps"], start=1):
        ext = [hexf(x) for x in st["ext"]]
        dt_arg = None if st["dtArg"] is None else hexf(st["dtArg"])
        dt = sim.step(dt_arg, ext)
        if sponge and dt > 0:
            sim.sponge(dt, hexf(sponge["etaRef"]))
        want_dt = hexf(st["dt"])
        if not st["finite"]:
            if sim.finite():
                problems.append(f"step {n}: reference went non-finite, solver did not")
            break
        if exact:
            if dt != want_dt:
                problems.append(f"step {n}: dt {dt!r} != {want_dt!r}")
                break
            mf = sim.mass_floored
            if mf != hexf(st["mf"]):
                problems.append(f"step {n}: massFloored {mf!r} != {hexf(st['mf'])!r}")
                break
            fp = fingerprint(sim)
            if fp != int(st["hash"], 16):
                problems.append(f"step {n}: fingerprint {fp:016x} != {st['hash']}")
                break
        else:
            if not (abs(dt - want_dt) <= rtol * abs(want_dt)):
                problems.append(f"step {n}: dt {dt!r} vs {want_dt!r}")
                break
            cp = fx["checkpoints"].get(str(n))
            if cp:
                h, hu, hv = sim.state()[1:]
                for name, got in (("h", h), ("hu", hu), ("hv", hv)):
                    want = arr(cp[name])
                    scale = max(np.max(np.abs(want)), 1e-300)
                    err = np.max(np.abs(got - want)) / scale
                    if not err <= rtol:
                        problems.append(f"step {n}: {name} rel err {err:.3e} > {rtol:g}")
                wmask = arr(cp["h"]) <= md
                if not np.array_equal(h <= md, wmask):
                    problems.append(f"step {n}: wet/dry mask differs in {int(np.sum((h <= md) != wmask))} cells")
                if problems:
                    break
        if on_step:
            on_step(n, sim)
    return problems
