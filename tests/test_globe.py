"""The globe harness (water-simulator_amd/globe.py, restating src/globe.mjs's
measureModePeriod and polarTimestepTax) on the host: the polar timestep tax off
the metric, the measurement's refusals, and its set-up against closed forms."""
import math

import numpy as np
import pytest

import globe
import swe_amd


def test_polar_timestep_tax_off_the_metric():
    # the reference's check (globe.mjs:2891-2895): ratio > 5 on the 3-degree sphere
    rows = swe_amd.geometry_spherical(120, 60)
    lat, _ = swe_amd.sphere_lonlat(120, 60)
    tax = globe.polar_timestep_tax(rows, lat)
    assert tax["ratio"] > 5
    assert abs(abs(tax["pole_lat"]) - 88.5) < 1e-9 and abs(abs(tax["eq_lat"]) - 1.5) < 1e-9
    # independent estimate from the cell sizes: dx = R dlam cos(lat), dy = R dphi, and
    # at the last row the pole-side face is a point, so its y divisor is half a cell
    d = math.radians(3.0)
    est = (1 / (math.cos(math.radians(88.5)) * d) + 1 / (d / 2)) / (1 / (math.cos(math.radians(1.5)) * d) + 1 / d)
    assert abs(tax["ratio"] / est - 1) < 0.02, (tax["ratio"], est)
    # a latitude cap at +-80 removes most of the tax
    rows_c = swe_amd.geometry_spherical(120, 53, lat0=-79.5, lat1=79.5)
    lat_c, _ = swe_amd.sphere_lonlat(120, 53, -79.5, 79.5)
    capped = globe.polar_timestep_tax(rows_c, lat_c)
    assert capped["ratio"] < tax["ratio"] / 4 and abs(abs(capped["pole_lat"]) - 78.0) < 1e-9


def test_measurement_refuses_its_own_g_and_rigid_modes():
    with pytest.raises(ValueError, match="g_ref"):
        globe.measure_mode_period(n=2, deg=12)
    with pytest.raises(ValueError):
        globe.legendre_p(1, 0.3)


def test_mode_problem_set_up():
    m = globe.mode_problem(n=2, deg=6)
    assert (m["nx"], m["ny"]) == (60, 30)
    # the rows' areas sum to the sphere's, and the projection norm is the area
    # integral of P2^2 = 4 pi R^2 / 5 to the midpoint rule's accuracy
    R = 6371e3
    area = m["nx"] * np.sum(m["rows"]["area"][2:-2])
    assert abs(area / (4 * math.pi * R * R) - 1) < 1e-12
    assert abs(m["norm"] / (4 * math.pi * R * R / 5) - 1) < 2e-3
    # the initial state projects onto itself with amplitude ~1 (the mode's own)
    c0 = globe.projection(m["coef"], m["b"], m["h"], m["nx"], m["ny"]) / m["norm"]
    assert abs(c0 - 1) < 1e-3
    # P3 and P2 are orthogonal on the grid to the midpoint rule's accuracy
    m3 = globe.mode_problem(n=3, deg=6)
    cross = globe.projection(m["coef"], m3["b"], m3["h"], m["nx"], m["ny"]) / m["norm"]
    assert abs(cross) < 1e-3
