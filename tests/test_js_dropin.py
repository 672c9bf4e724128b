"""The Node drop-in module (water-simulator_amd/js/swe.mjs over the N-API addon).

CPU: its host pieces (hllc export, metric rows, sponge strengths) against the
reference's recorded values.  GPU: every golden scenario replayed through the
drop-in end to end, and the restated verification suites run against it.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
needs_node = pytest.mark.skipif(NODE is None, reason="node is not installed")
ADDON = os.path.join(ROOT, "water-simulator_amd", "napi", "swe_napi.node")


def run_node(script, *args, timeout=600, env=None):
    print(f"[node] {script} {' '.join(args)}", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, script), *args], cwd=ROOT, capture_output=True, text=True,
                       timeout=timeout, env=env)
    sys.stdout.write(r.stdout[-6000:])
    sys.stderr.write(r.stderr[-3000:])
    return r.returncode


@needs_node
def test_js_host_pieces():
    if not os.path.exists(ADDON):
        pytest.skip("addon not built")
    assert run_node("tests/js/cpu_checks.mjs") == 0


@needs_node
def test_addon_refuses_arrays_of_the_wrong_length(tmp_path):
    """Every addon binding that passes a typed array to the C library checks its
    length first (W*H fields, H rows, samples x gauges series, one tide value per
    constituent) and throws a RangeError, never calling the library with it.  Run
    on the CPU against a stand-in libswe_amd.so (tests/napi_mock/swe_mock.c) that
    the addon loads through LD_LIBRARY_PATH."""
    if not os.path.exists(ADDON):
        pytest.skip("addon not built")
    lib = tmp_path / "libswe_amd.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-o", str(lib),
                    os.path.join(ROOT, "tests", "napi_mock", "swe_mock.c")], check=True)
    env = dict(os.environ, LD_LIBRARY_PATH=f"{tmp_path}:{os.environ.get('LD_LIBRARY_PATH', '')}")
    assert run_node("tests/js/napi_lengths_cpu.mjs", env=env, timeout=120) == 0


@needs_node
@pytest.mark.gpu
def test_js_golden_replay_on_gpu():
    assert run_node("tests/js/golden_replay.mjs") == 0


@needs_node
@pytest.mark.gpu
def test_js_restated_verification_suites_on_gpu():
    """Lake at rest, mass, Ritter, celerity, Thacker, Merian, positivity, Manning
    spin-down, Coriolis map, CFL, run-up and Flather, through the drop-in."""
    assert run_node("tests/js/verify_gpu.mjs", timeout=900) == 0


@needs_node
@pytest.mark.gpu
def test_js_adjoint_and_tide_on_gpu():
    assert run_node("tests/js/adjoint_gpu.mjs") == 0


@needs_node
@pytest.mark.gpu
def test_js_device_diagnostics_and_recorded_runs_on_gpu():
    assert run_node("tests/js/record_gpu.mjs") == 0


@needs_node
@pytest.mark.gpu
def test_js_gridlands_import_gates_on_gpu():
    """tools/verify-gridlands.mjs's gates through js/gridlands.mjs and the drop-in."""
    assert run_node("tests/js/gridlands_gpu.mjs") == 0


@needs_node
@pytest.mark.gpu
def test_js_globe_harness_on_gpu():
    """src/globe.mjs's measureModePeriod / polarTimestepTax checks through js/globe.mjs."""
    assert run_node("tests/js/globe_gpu.mjs") == 0


@needs_node
@pytest.mark.gpu
def test_js_reference_recorded_still_water_figures_on_gpu():
    """README.md:50-55's lake-at-rest table and swe.mjs:821-823's drift, reproduced exactly."""
    assert run_node("tests/js/reference_numbers_gpu.mjs") == 0


@needs_node
@pytest.mark.gpu
def test_js_verify_sections_to_the_printed_digit_on_gpu():
    """tools/verify.mjs sections 3-8 (Ritter, celerity, Thacker, Merian, self-convergence,
    dissipation) through the drop-in: every figure equal to the one the reference prints."""
    assert run_node("tests/js/verify_sections_gpu.mjs") == 0


@needs_node
@pytest.mark.gpu
def test_js_verify_physics_sections_to_the_printed_digit_on_gpu():
    """tools/verify-physics.mjs sections 1-10 (Coriolis, Manning, wetting and drying,
    positivity, vel and run-up, isotropy, the contact wave, gravity, the CFL step, hllc's wave
    speeds, the open boundaries) through the drop-in: every printed figure equal."""
    assert run_node("tests/js/verify_physics_sections_gpu.mjs") == 0


@needs_node
@pytest.mark.gpu
def test_js_verify_tide_sections_to_the_printed_digit_on_gpu():
    """tools/verify-tide.mjs sections 3, 3b and 3c (Flather absorption against a wall, the
    oblique-incidence and finite-amplitude tables) through the drop-in: every printed figure equal."""
    assert run_node("tests/js/verify_tide_sections_gpu.mjs") == 0
