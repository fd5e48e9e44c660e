"""CPU: every pinned oracle rule is load-bearing.  The oracle is rebuilt
thirteen times with one rule broken each (oracle.cpp's ORC_MUTANT switches:
sameSurface without its time test, ternaryEqual needing equal signs, the leaf scan
without the cocboundst0[0] typo or ignoring time changes, fetch2R ignoring the
quadrant mask, no tiny-surface merge, emergency passes keeping the
dispersion, no nearest-sample fallback, output times without the jitter,
the GPU-variant 2R surface split at 3 points, shadow queries without the
density scaling, the shadow bias added instead of subtracted, and the shadow
fallback without its light test), and the pin tests must fail on each, in the
named test; on the unbroken oracle they pass."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
PINS = os.path.join(ROOT, "tests", "test_oracle_pins.py")
SHADOW_PINS = os.path.join(ROOT, "tests", "test_oracle_shadow.py")

# mutant -> a pin test that must catch it
CATCHES = {
    1: "test_same_surface_crossing_in_time",
    2: "test_same_surface_one_sided_difference",
    3: "test_scan_cocbounds_typo_decides",
    4: "test_scan_time_split",
    5: "test_no_fetch2r_when_first_surface_has_all_quadrants",
    6: "test_tiny_front_surface_merges_into_back",
    7: "test_emergency_pass_gives_nearest_sample",
    8: "test_emergency_pass_gives_nearest_sample",
    9: "test_output_pattern_composition",
    10: "test_refgpu_2r_surface_needs_four_points",
    11: "test_density_scales_the_dispersion",
    12: "test_bias_keeps_a_query_just_behind_lit",
    13: "test_nearest_fallback_takes_the_light_test",
}


@pytest.fixture(scope="module")
def mutants():
    subprocess.run(["make", "-s", "-C", ORACLE, "-j8", "mutants"], check=True)
    return {k: os.path.join(ORACLE, "_build", "liboracle_m%d.so" % k) for k in CATCHES}


def _run_pins(lib, select=None):
    env = dict(os.environ, TLR_ORACLE_LIB=lib)
    cmd = [sys.executable, "-m", "pytest", PINS, SHADOW_PINS, "-q", "-p", "no:cacheprovider"]
    if select:
        cmd += ["-k", select]
    return subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=300)


def test_pins_pass_on_the_oracle():
    r = _run_pins(os.path.join(ORACLE, "_build", "liboracle.so"))
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("k", sorted(CATCHES))
def test_pins_catch_mutant(mutants, k):
    r = _run_pins(mutants[k], CATCHES[k])
    print(r.stdout[-600:])
    assert r.returncode == 1 and "failed" in r.stdout, "mutant %d passes %s" % (k, CATCHES[k])
