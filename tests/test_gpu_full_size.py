"""The benchmark workloads at their full sizes on the GPU, against the CPU oracle:
  the Cornell box, 256x256, 16 spp, max depth 8;
  the 1.3M-triangle displaced sphere under the 512x256 sky map, 1024x1024, 64 spp, max depth 6;
  the heterogeneous-medium Cornell box (256^3 fBm density), 512x512, 128 spp, max depth 16.
Every GPU render is the whole workload (n_gpu seeds). The oracle renders a spread of the image's
32x32 blocks (the reference's jobs, seeded (blockX, blockY)) at the same spp with n_cpu seed offsets.
The means of the 8x8 pixel tiles inside those blocks are compared with a two-sample z-test (each
side's standard error from its own renders, plus a relative floor of 1e-4): max |z| < 6 and mean
z^2 < 2 over tiles and channels (expected ~1.1 for the t-distributed statistic). The GPU draws
counter-based streams and the CPU per-block PCG32 streams, so only expectations can agree."""
import os

import numpy as np
import pytest

from raytracer_amd import rtg, scenes
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


def _tiles(img, mask, t=8):
    H, W = mask.shape
    m = mask[: H // t * t, : W // t * t].reshape(H // t, t, W // t, t).all(axis=(1, 3))
    x = img[: H // t * t, : W // t * t].reshape(H // t, t, W // t, t, 3).mean(axis=(1, 3))
    return x[m]


def _full_size_z(ctx, sd, n_gpu, n_cpu, stride, label):
    cam, rp = sd.camera, sd.render
    W, H = int(cam["width"]), int(cam["height"])
    spp = rp["spp"]
    if rp.get("sampler", rtg.SAMPLER_STRATIFIED) == rtg.SAMPLER_STRATIFIED:
        spp = int(np.ceil(np.sqrt(spp))) ** 2  # a non-square stratified count rounds up to a square
    g = ctx.upload(sd)
    print(f"[full size] {label}: {W}x{H} {rp['spp']} spp depth {rp['max_depth']}, gpu {n_gpu} renders, "
          f"cpu {n_cpu} renders of every {stride}th block", flush=True)
    gpu = []
    for s in range(n_gpu):
        out, st = g.render(cam, dict(rp, seed=101 + s))
        assert st.samples == W * H * spp and st.nan_samples == 0
        gpu.append(rtg.resolve(out).astype(np.float64))
    print(f"  gpu {st.milliseconds:.1f} ms per render, {st.samples / st.milliseconds / 1e3:.0f} M samples/s",
          flush=True)
    o = O.OracleScene(sd)
    nblocks = ((W + 31) // 32) * ((H + 31) // 32)
    cpu, mask = [], None
    for s in range(n_cpu):
        rgb, w, stats, nb = o.render(camera=cam, threads=min(16, os.cpu_count()), blocks=(0, nblocks, stride),
                                     seed_offset=s)
        mask = w > 0 if mask is None else mask & (w > 0)
        cpu.append(rgb / np.maximum(w[..., None], 1e-30))
    gt = np.stack([_tiles(x, mask) for x in gpu])
    ct = np.stack([_tiles(x, mask) for x in cpu])
    assert gt.shape[1] >= 64, "too few tiles compared"
    mg, mc = gt.mean(0), ct.mean(0)
    se2 = gt.var(0, ddof=1) / n_gpu + ct.var(0, ddof=1) / n_cpu + (1e-4 * (np.abs(mc) + 1e-3)) ** 2
    z = (mg - mc) / np.sqrt(se2)
    print(f"  {gt.shape[1]} tiles: max|z| {np.abs(z).max():.2f} mean z^2 {np.mean(z * z):.2f} "
          f"mean gpu {mg.mean():.5f} cpu {mc.mean():.5f}", flush=True)
    assert np.abs(z).max() < 6.0 and np.mean(z * z) < 2.0
    g.release()


def test_cornell_full_size(ctx):
    _full_size_z(ctx, scenes.cornell(), n_gpu=16, n_cpu=32, stride=3, label="cornell")


def test_displaced_sphere_full_size(ctx):
    _full_size_z(ctx, scenes.displaced_sphere(), n_gpu=16, n_cpu=32, stride=61, label="displaced sphere")


def test_vol_cornell_full_size(ctx):
    sd = scenes.vol_cornell()
    assert sd.render["integrator"] == rtg.INTEGRATOR_VOLPATH and sd.render["max_depth"] == 16
    _full_size_z(ctx, sd, n_gpu=16, n_cpu=16, stride=30, label="volumetric cornell")
