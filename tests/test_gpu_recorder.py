"""The device gauge recorder of swe_advance against the oracle.

Bar: bit-exact.  A recorded device-resident run samples eta = b + h at the
gauges exactly where a host loop over the oracle's steps does: every k-th
step, or on the sample-time schedule of tools/waves.mjs:211-232 (steps
shortened to min(maxDt, max(1e-6, min(T, next) - t)), a sample after the
first step ending at or after each sample time from the analysis start on).
The recorder must not change the run itself.
"""
import sys

import numpy as np
import pytest

import swe_amd
from gpu_util import bits, pair

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gauges(nx, ny, n, seed):
    rng = np.random.default_rng(seed)
    W = nx + 4
    i = rng.integers(2, nx + 2, n)
    j = rng.integers(2, ny + 2, n)
    return (j * W + i).astype(np.int32)


def jmin(a, b):
    return a if a < b else b      # Math.min for the non-NaN values met here


def jmax(a, b):
    return a if a > b else b


@pytest.mark.parametrize("graph", [False, True])
def test_recorder_every_k_steps(graph):
    say(f"[gpu] recorder every 3 steps, graph={graph}")
    nx, ny = 130, 70
    bc = ("reflect", "outflow", "periodic", "periodic")
    g, o = pair(nx, ny, 31, bc=bc, manning=0.02)
    ref, _ = pair(nx, ny, 31, bc=bc, manning=0.02)
    cells = gauges(nx, ny, 37, 3)
    g.set_recorder(cells, every=3, capacity=20)
    g.use_graph(graph)
    assert g.advance(31) == 31
    assert ref.advance(31) == 31
    times, eta = g.record()
    want_t, want = [], []
    b = o.b.copy()
    for n in range(1, 32):
        o.step(None)
        if n % 3 == 0:
            want_t.append(o.t)
            want.append(b[cells] + o.h[cells])
    assert np.array_equal(bits(times), bits(np.array(want_t)))
    assert np.array_equal(bits(eta), bits(np.array(want)))
    # the run itself is the unrecorded run's, bit for bit
    for x, y in zip(g.state(), ref.state()):
        assert np.array_equal(bits(x), bits(y))


@pytest.mark.parametrize("graph", [False, True])
def test_recorder_sample_times_follow_waves_schedule(graph):
    say(f"[gpu] recorder on sample times, graph={graph}")
    nx, ny = 96, 48
    bc = ("outflow", "reflect", "periodic", "periodic")
    g, o = pair(nx, ny, 12, bc=bc, manning=0.025)
    cells = gauges(nx, ny, 50, 8)
    T0, dt_rec, T = 0.9, 0.125, 3.0
    g.set_recorder(cells, every=0, t0=T0, dt_rec=dt_rec, capacity=64)
    g.use_graph(graph)
    taken = g.advance(100000, swe_amd.DT_SAMPLED, t_end=T)
    times, eta = g.record()
    # tools/waves.mjs:221-228 on the oracle
    nxt, want_t, want, steps = T0, [], [], 0
    b = o.b.copy()
    while o.t < T:
        o.step(jmin(o.max_dt(), jmax(1e-6, min(T, nxt) - o.t)))
        steps += 1
        if o.t >= nxt:
            if o.t >= T0:
                want_t.append(o.t)
                want.append(b[cells] + o.h[cells])
            nxt += dt_rec
    say(f"  {steps} steps, {len(want_t)} samples, last t {o.t!r}")
    assert taken == steps
    assert g.t == o.t
    assert len(want_t) >= 16
    assert np.array_equal(bits(times), bits(np.array(want_t)))
    assert np.array_equal(bits(eta), bits(np.array(want)))
    _, h, hu, hv = g.state()
    for x, y in ((h, o.h), (hu, o.hu), (hv, o.hv)):
        assert np.array_equal(bits(x), bits(y))


def test_recorder_limits_and_errors():
    say("[gpu] recorder limits")
    nx, ny = 40, 30
    g, _ = pair(nx, ny, 2)
    cells = gauges(nx, ny, 4, 1)
    g.set_recorder(cells, every=1, capacity=3)
    assert g.advance(5) == 5
    with pytest.raises(swe_amd.SweError, match="capacity"):
        g.record()
    with pytest.raises(swe_amd.SweError, match="out of range"):
        g.set_recorder(np.array([(nx + 4) * (ny + 4)], np.int32), every=1)
    with pytest.raises(swe_amd.SweError, match="sample times"):
        g.set_recorder(cells, every=0, dt_rec=0.0)
    g.set_recorder([], every=0)
    with pytest.raises(swe_amd.SweError, match="SAMPLED"):
        g.advance(3, swe_amd.DT_SAMPLED, t_end=1.0)
    # re-armed: the samples start again from zero
    g.set_recorder(cells, every=2, capacity=8)
    assert g.advance(4) == 4
    times, eta = g.record()
    assert eta.shape == (2, 4) and times[0] < times[1]
