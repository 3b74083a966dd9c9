"""Allocation-free trace (rtg_trace_device_noalloc, through the C ABI): for caller-owned device
buffers it returns exactly what rtg_trace returns (primitive ids, t, barycentrics, hit point and
normal bit for bit; any-hit flags), on two calls in a row, the second with fewer rays (the shared
scratch keeps the first call's contents, so its counters must be cleared per call); it spills
into the static scratch on the 1.3M-triangle scene and rejects nothing there (stack bound 35 of
48); a null ray buffer is refused."""
import ctypes as C

import numpy as np
import pytest
# torch (device buffers) is imported before librtg creates a context: torch ships its own HIP
# runtime, which then serves both
import torch

from raytracer_amd import rtg, scenes
from gpu_util import random_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


def _noalloc(ctx, scene, rays, any_hit=False):
    n = rays.shape[0]
    dev = torch.device("cuda")
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(dev)
    prim = torch.full((n,), -7, dtype=torch.int32, device=dev)
    t = torch.zeros(n, dtype=torch.float32, device=dev)
    uv = torch.zeros(2 * n, dtype=torch.float32, device=dev)
    p = torch.zeros(3 * n, dtype=torch.float32, device=dev)
    nrm = torch.zeros(3 * n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    ptr = lambda x: C.c_void_p(x.data_ptr())
    ctx.check(ctx.lib.rtg_trace_device_noalloc(ctx.h, scene.h, ptr(r), n, int(any_hit), ptr(prim), ptr(t), ptr(uv),
                                               ptr(p), ptr(nrm)))
    return (prim.cpu().numpy(), t.cpu().numpy(), uv.cpu().numpy().reshape(n, 2), p.cpu().numpy().reshape(n, 3),
            nrm.cpu().numpy().reshape(n, 3))


def _same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("level", [8, 5])
def test_noalloc_trace_equals_trace_on_two_calls(ctx, level):
    sd = scenes.displaced_sphere(level=level)
    g = ctx.upload(sd)
    assert g.info()["traversal_stack"] <= 48
    rng = np.random.default_rng(30 + level)
    big = random_rays(rng, 200000, np.array([0, 1.1, 0]), 1.3)
    small = random_rays(rng, 3000, np.array([0, 1.1, 0]), 1.3)
    for rays in (big, small):  # the second call has fewer rays than the first
        ref = g.trace(rays)
        ref_spills = ctx.trace_spills()
        got = _noalloc(ctx, g, rays)
        for a, b in zip(got, ref):
            assert _same_bits(a, b)
        if level == 8 and rays is big:
            assert ref_spills > 0 and ctx.trace_spills() > 0  # the spill area in the scratch was used
        assert np.array_equal(_noalloc(ctx, g, rays, any_hit=True)[0] >= 0, g.trace(rays, any_hit=True)[0] >= 0)
    g.release()


def test_noalloc_trace_refuses_a_null_ray_buffer(ctx):
    g = ctx.upload(scenes.cornell())
    rc = ctx.lib.rtg_trace_device_noalloc(ctx.h, g.h, None, 10, 0, None, None, None, None, None)
    assert rc == -1  # RTG_ERR_INVALID
    g.release()
