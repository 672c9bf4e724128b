"""Batched ensembles on the GPU: every basin of a batch must be bit-identical to
a plain single-basin run of the same basin (same dt sequence, same state), with
and without graph replay, whatever the task shape the autotuner picks."""
import sys

import numpy as np
import pytest

import swe_amd
import workloads
from gpu_util import bits

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def build(specs, kind):
    if kind == "batch":
        B = swe_amd.Batch([dict(nx=s["nx"], ny=s["ny"], manning=s["manning"], cfl=s["cfl"],
                                rows=swe_amd.geometry_cartesian(s["nx"], s["ny"], s["dx"])) for s in specs])
        for i, s in enumerate(specs):
            B.load_state(i, *s["state"])
            B.set_bc(i, s["bc"])
            B.set_harmonic(i, 0, amp=[s["amp"]], omega=[s["omega"]], phase=[0.0],
                           uamp=[s["uamp"]], uomega=[s["omega"]], uphase=[0.0])
        return B
    sims = []
    for s in specs:
        g = swe_amd.Solver(s["nx"], s["ny"], swe_amd.geometry_cartesian(s["nx"], s["ny"], s["dx"]),
                           manning=s["manning"], cfl=s["cfl"])
        g.load_state(*s["state"])
        g.set_bc(s["bc"])
        g.set_harmonic(0, amp=[s["amp"]], omega=[s["omega"]], phase=[0.0],
                       uamp=[s["uamp"]], uomega=[s["omega"]], uphase=[0.0])
        sims.append(g)
    return sims


@pytest.mark.parametrize("graph", [True, False])
def test_batch_basins_bit_identical_to_plain_runs(graph):
    specs = workloads.ensemble(nb=9, lo=6, hi=180, seed=17)
    say(f"[gpu] batch vs plain, graph={graph}, shapes {[(s['nx'], s['ny']) for s in specs]}")
    B = build(specs, "batch")
    B.use_graph(graph)
    B.set_chunk_rows(16)
    plain = build(specs, "plain")
    assert B.advance(12) == 12
    for g in plain:
        assert g.advance(12) == 12
    for i, g in enumerate(plain):
        for a, b in zip(B.state(i), g.state()):
            assert np.array_equal(bits(a), bits(b)), i
        assert B.get_scalar(i, swe_amd.S_T) == g.t


def test_batch_autotune_keeps_state_and_results():
    specs = workloads.ensemble(nb=5, lo=40, hi=300, seed=23)
    say("[gpu] batch autotune")
    B = build(specs, "batch")
    before = [B.state(i) for i in range(len(specs))]
    rows, waves = B.autotune(steps=1)
    assert rows >= 8 and waves >= 1
    for i in range(len(specs)):
        for a, b in zip(B.state(i), before[i]):
            assert np.array_equal(bits(a), bits(b))
        assert B.get_scalar(i, swe_amd.S_STEPS) == 0
    plain = build(specs, "plain")
    B.advance(6)
    for i, g in enumerate(plain):
        g.advance(6)
        for a, b in zip(B.state(i), g.state()):
            assert np.array_equal(bits(a), bits(b)), i
