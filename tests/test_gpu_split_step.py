"""The step split around a host ghost fill (swe_step_begin / swe_step_end) and the
residual a step leaves, on the GPU, against the oracle (the reference's step
restated on the CPU, src/swe.mjs:754-814).

* With every side SWE_BC_STORED and the ghost ring filled on the host before
  each stage (the way the Node drop-in runs a caller's own boundary function,
  swe.mjs:289-303), the GPU reproduces the oracle's step bit for bit.
* With device boundary kinds the split step equals swe_step bit for bit.
* After a step, SWE_F_RH/RHU/RHV/THETA hold stage 2's residual and theta, what
  the reference leaves in sim.rh and sim.theta (swe.mjs:798-806), bit for bit.
"""
import numpy as np
import pytest

import oracle
import swe_amd
from gpu_util import bits, pair, random_basin

pytestmark = pytest.mark.gpu

BC = ("reflect", "outflow", "periodic", "periodic")


def host_fill(filler, b, h, hu, hv):
    """The ghost ring of (b, h, hu, hv) as the oracle's applyBC writes it."""
    filler.load_state(b, h, hu, hv)
    filler.apply_bc()
    return filler.state()


def test_split_step_with_host_ghost_fill_matches_the_oracle():
    nx, ny = 90, 56
    g, o = pair(nx, ny, 31, bc=BC, manning=0.025, coriolis=1e-4)
    rows = swe_amd.geometry_cartesian(nx, ny, 5.0, coriolis=1e-4)
    filler = oracle.OracleSim(nx, ny, rows, manning=0.025, coriolis=1e-4, dx=5.0, dy=5.0)
    filler.set_bc(BC)
    g.set_bc(("stored",) * 4)
    for n in range(25):
        st = host_fill(filler, *g.state())
        g.load_state(*st)
        dt = g.step_begin(None)
        do = o.step(None)
        assert dt == do and dt > 0, (n, dt, do)
        b = g.get_field(swe_amd.F_B)
        st = host_fill(filler, b, *(g.get_field(f) for f in (swe_amd.F_H, swe_amd.F_HU, swe_amd.F_HV)))
        g.load_state(*st)
        g.step_end(want_theta=True, want_rh=True)
        _, h, hu, hv = g.state()
        for name, x, y in (("h", h, o.h), ("hu", hu, o.hu), ("hv", hv, o.hv)):
            assert np.array_equal(bits(x), bits(y)), (n, name)
        assert g.mass_floored == o.mass_floored
    for f, of in ((swe_amd.F_RH, oracle.RH), (swe_amd.F_RHU, oracle.RHU), (swe_amd.F_RHV, oracle.RHV),
                  (swe_amd.F_THETA, oracle.THETA)):
        assert np.array_equal(bits(g.get_field(f)), bits(o.field(of))), f


def test_split_step_with_device_kinds_equals_swe_step():
    nx, ny = 77, 63
    bc = ("flather", "reflect", "reflect", "outflow")
    a, _ = pair(nx, ny, 32, bc=bc, manning=0.02)
    b, _ = pair(nx, ny, 32, bc=bc, manning=0.02)
    ext = [0.1, 0.02, 0, 0, 0, 0, 0, 0]
    for n in range(20):
        da = a.step(None, ext)
        db = b.step_begin(None, ext)
        b.step_end()
        assert da == db
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(bits(x), bits(y))


def test_residual_left_by_a_step_matches_the_oracle():
    nx, ny = 64, 70
    g, o = pair(nx, ny, 33, bc=("reflect", "reflect", "periodic", "periodic"), manning=0.025, coriolis=1e-4)
    for n in range(12):
        dg = g.step(None, None, want_theta=True, want_rh=True)
        do = o.step(None)
        assert dg == do
        for f, of in ((swe_amd.F_RH, oracle.RH), (swe_amd.F_RHU, oracle.RHU), (swe_amd.F_RHV, oracle.RHV),
                      (swe_amd.F_THETA, oracle.THETA)):
            assert np.array_equal(bits(g.get_field(f)), bits(o.field(of))), (n, f)
    # keep_residual: swe_advance leaves the last step's residual too
    g.keep_residual(True)
    g.advance(5)
    o.run(5)
    for f, of in ((swe_amd.F_RH, oracle.RH), (swe_amd.F_THETA, oracle.THETA)):
        assert np.array_equal(bits(g.get_field(f)), bits(o.field(of))), f


def test_ghost_ring_download_and_mid_step_guard():
    nx, ny = 40, 30
    g, o = pair(nx, ny, 34, bc=("reflect", "outflow", "reflect", "periodic"))
    g.step(None)
    full = g.get_field(swe_amd.F_B)
    ring = np.full(g.N, -7.0)
    g.get_ghosts(swe_amd.F_B, ring)
    m = np.ones((ny + 4, nx + 4), bool)
    m[2:-2, 2:-2] = False
    m = m.ravel()
    assert np.array_equal(ring[m], full[m]) and np.all(ring[~m] == -7.0)
    g.set_bc(("stored",) * 4)
    assert g.step_begin(None) > 0
    with pytest.raises(swe_amd.SweError):
        g.advance(1)
    with pytest.raises(swe_amd.SweError):
        g.step(None)
    g.step_end()
    g.set_bc(("reflect",) * 4)
    assert g.step(None) > 0
