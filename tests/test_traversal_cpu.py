"""The BVH builder's 4-wide nodes and the device traversal code, checked on the CPU.

tests/cpu_trav/trav_check.cpp compiles rtg_bvh.cpp and the traversal functions of rtg_device.h
(trav_begin, trav_step, the LDS/global stack helpers) with g++ through a CPU stand-in for the device
intrinsics, under AddressSanitizer and UBSan, then traces random rays (zero direction components,
short and long intervals, adaptive epsilon) through random triangle/sphere scenes and compares
closest and any hits with brute force over all primitives using the same intersection code. It also
checks that every primitive sits in exactly one reachable leaf and that the deepest stack a ray
reached stays within the builder's bound (which sizes the GPU stack)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "raytracer_amd", "csrc")
BUILD = os.path.join(HERE, "_build")


@pytest.fixture(scope="module")
def trav_check():
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, f"trav_check.{os.getpid()}")  # per process: parallel workers build their own
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-include", os.path.join(HERE, "cpu_trav", "device_cpu.h"), "-I" + os.path.join(HERE, "mock_hip"),
           "-I" + CSRC, os.path.join(HERE, "cpu_trav", "trav_check.cpp"), os.path.join(CSRC, "rtg_bvh.cpp"),
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.mark.parametrize("seed,ntri,nsph,nrays", [(1, 5000, 0, 20000), (2, 3000, 200, 20000), (3, 1, 0, 2000),
                                                  (4, 0, 5, 2000), (5, 40000, 0, 6000), (6, 200, 60, 20000)])
def test_traversal_matches_brute_force(trav_check, seed, ntri, nsph, nrays):
    r = subprocess.run([trav_check, str(seed), str(ntri), str(nsph), str(nrays)], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), (r.stdout[-2000:], r.stderr[-3000:])


def test_stack_overflow_is_flagged(trav_check):
    """A stack smaller than what the rays need (forced to 16 entries on a 40k-triangle scene whose
    rays reach depth 21) makes pushes find it full: the traversal raises its overflow
    flag (which the library turns into a render error) instead of dropping subtrees silently."""
    r = subprocess.run([trav_check, "5", "40000", "0", "6000", "16"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "OVERFLOW FLAGGED" in r.stdout, (r.stdout[-2000:], r.stderr[-3000:])


def test_sampler_strata_arithmetic():
    """The device sampler's stratum arithmetic and permutation (rtg_device.h in_stratum, permute),
    compiled for the host: every stratified value stays inside its cell in exact arithmetic (also
    for jitters within an ulp of 1, where the float sum cell + jitter rounds up to the next cell),
    and the permutation is a bijection, for sample counts up to 4225."""
    exe = os.path.join(BUILD, f"sampler_check.{os.getpid()}")
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-include", os.path.join(HERE, "cpu_trav", "device_cpu.h"),
           "-I" + os.path.join(HERE, "mock_hip"), "-I" + CSRC, os.path.join(HERE, "cpu_trav", "sampler_check.cpp"),
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-2000:]


def test_guided_cdf_search_equals_full_search():
    """The environment light's CDF search with cut points (rtg_device.h cdf_sample_guided, cut
    points from rtg_internal.h cdf_guide) returns the full binary search's index for random draws,
    every CDF value and every cut value with their float neighbours, on CDFs with zero runs, one
    dominant entry and heavy skew; the cut points equal their directly counted definition."""
    exe = os.path.join(BUILD, f"cdf_check.{os.getpid()}")
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-include", os.path.join(HERE, "cpu_trav", "device_cpu.h"),
           "-I" + os.path.join(HERE, "mock_hip"), "-I" + CSRC, os.path.join(HERE, "cpu_trav", "cdf_check.cpp"),
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-2000:]
