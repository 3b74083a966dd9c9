"""Serving-loop host logic on the CPU: librtg's host sources (serving loop, job table, scene
upload) built with g++ against a HIP stand-in (tests/mock_hip) and driven through the same
ctypes binding and scenarios as tests/test_gpu_serving.py, under glibc's heap checking
(MALLOC_CHECK_=3). Regression test for a use-after-free in rtg_server_destroy (the request-table
mutex was unlocked after the server object holding it had been deleted), which glibc reports as
"free(): corrupted unsorted chunks"."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "raytracer_amd", "csrc")
MOCK = os.path.join(ROOT, "tests", "mock_hip")
BUILD = os.path.join(ROOT, "tests", "_build")

RUNNER = r"""
import sys
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
from raytracer_amd import rtg, scenes
rtg.LIB_PATH = {lib!r}
import test_gpu_serving as T
ctx = rtg.Context(0)
ctx.set_pool(1 << 16)
res = {{}}
for name, sd in (("cornell", scenes.cornell()), ("glossy", scenes.glossy_cornell()),
                 ("displaced", scenes.displaced_sphere(level=3)), ("vol", scenes.vol_cornell(res=16))):
    res[name] = (sd, ctx.upload(sd))
for fn in [T.test_same_alone_and_in_mix, T.test_render_on_a_serving_context_joins_the_mix,
           T.test_volumetric_requests_share_the_pool, T.test_budget_admission_serialises_films,
           T.test_cancel_drains_and_others_finish, T.test_submit_from_another_thread_while_stepping,
           lambda c, r: T.server_thread_and_wait(c, r, device_film=False)]:
    fn(ctx, res)
    print("PASS", getattr(fn, "__name__", "?"), flush=True)
ctx.close()
print("ALL OK", flush=True)
"""


def build_mock_lib():
    os.makedirs(BUILD, exist_ok=True)
    # per process: parallel test workers each build (and load) their own copy
    out = os.path.join(BUILD, f"librtg_hostmock.{os.getpid()}.so")
    srcs = [os.path.join(CSRC, f) for f in ("rtg_serve.cpp", "rtg_host.cpp", "rtg_guide_host.cpp", "rtg_bvh.cpp",
                                            "rtg_xml.cpp")]
    srcs.append(os.path.join(MOCK, "mock_kernels.cpp"))
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-pthread", "-Wl,--no-undefined", "-I" + MOCK,
           "-I" + CSRC] + srcs + ["-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def test_serving_host_logic_under_heap_checks():
    lib = build_mock_lib()
    code = RUNNER.format(root=ROOT, tests=os.path.join(ROOT, "tests"), lib=lib)
    env = dict(os.environ, MALLOC_CHECK_="3", MALLOC_PERTURB_="165")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout[-2000:])
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-3000:])
    assert "ALL OK" in r.stdout


GUIDED_RUNNER = r"""
import sys
import numpy as np
sys.path.insert(0, {root!r})
from raytracer_amd import rtg, scenes
rtg.LIB_PATH = {lib!r}
ctx = rtg.Context(0)
ctx.set_pool(1 << 22)  # larger than the 2^21-slot cap guided renders train on
sd = scenes.cornell(width=64, height=64)
g = ctx.upload(sd)
spp = 1024  # 4.2 M samples: the training phase claims part of them on the capped pool
out, st = g.render(sd.camera, dict(sd.render, spp=spp, seed=3, guiding=1))
n = 64 * 64 * spp
assert st.samples == n, st.samples
# the stand-in adds (1, sample index, 0.5, 1) to pixel s mod (64*64) for every finished sample s
assert np.array_equal(out[..., 3].ravel(), np.full(64 * 64, float(spp), np.float32)), "sample counts per pixel"
assert np.array_equal(out[..., 1].ravel(), np.full(64 * 64, spp * (spp - 1) / 2.0, np.float32)), "sample sets"
ctx.close()
print("ALL OK", flush=True)
"""


def test_guided_render_pool_switch_host_logic():
    """A learned-guiding render on a context whose pool is larger than the guided cap trains on the
    capped pool, stops its claims there (the job's claim limit set to the claim counter), drains it
    and renders the rest on the full pool. Through the HIP stand-in, which finishes claimed samples
    and records each one in its pixel, every sample index must be rendered exactly once across the
    switch: per pixel exactly spp samples whose indices sum to spp (spp - 1) / 2."""
    lib = build_mock_lib()
    code = GUIDED_RUNNER.format(root=ROOT, lib=lib)
    env = dict(os.environ, MALLOC_CHECK_="3")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout[-2000:])
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-3000:])
    assert "ALL OK" in r.stdout
