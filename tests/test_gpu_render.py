"""GPU image parity with the CPU oracle (statistical: the GPU draws counter-based streams, the CPU
renderer draws per-block sequential PCG32 streams, so only the expectations can agree).

relMSE(x, r) = mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2)."""
import os

import numpy as np
import pytest

from raytracer_amd import rtg, scenes
import oracle_lib as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


def gpu_image(scene, cam, rp, **kw):
    out, st = scene.render(cam, dict(rp, **kw))
    return rtg.resolve(out).astype(np.float64), st


def cpu_image(osc, cam, spp, seed_offset, threads=None):
    rgb, w, st, _ = osc.render(camera=cam, spp=spp, threads=threads or min(16, os.cpu_count()),
                               seed_offset=seed_offset)
    return rgb / np.maximum(w[..., None], 1e-30)


def test_cornell_relmse_matches_cpu(ctx):
    """At 16 spp the GPU image's relMSE against the 65536-spp FP64 CPU reference matches the CPU
    renderer's own relMSE: ratio of the means over 8 independent renders each within [0.85, 1.18]
    (the CPU renders' relMSE has a relative spread of 11% per render, so the ratio's spread is ~5.5%)."""
    ref = np.load(os.path.join(GOLDEN, "cornell64_ref.npy")).astype(np.float64)
    sd = scenes.cornell()
    cam = dict(sd.camera, width=64, height=64)
    g = ctx.upload(sd)
    o = O.OracleScene(sd)
    print("[render] cornell 64x64 16spp, 8 GPU seeds vs 8 CPU seed offsets", flush=True)
    eg = [relmse(gpu_image(g, cam, sd.render, spp=16, seed=s)[0], ref) for s in range(1, 9)]
    ec = [relmse(cpu_image(o, cam, 16, s), ref) for s in range(8)]
    ratio = np.mean(eg) / np.mean(ec)
    print(f"  relMSE gpu {np.mean(eg):.5f} cpu {np.mean(ec):.5f} ratio {ratio:.3f}", flush=True)
    assert 0.85 < ratio < 1.18, (eg, ec)


def test_cornell_converges_to_reference(ctx):
    """Bias check: at 4096 spp the GPU image is within the noise of the 65536-spp reference."""
    ref = np.load(os.path.join(GOLDEN, "cornell64_ref.npy")).astype(np.float64)
    sd = scenes.cornell()
    cam = dict(sd.camera, width=64, height=64)
    g = ctx.upload(sd)
    print("[render] cornell 64x64 4096spp", flush=True)
    img, st = gpu_image(g, cam, sd.render, spp=4096, seed=7)
    assert st.nan_samples == 0
    e16 = relmse(gpu_image(g, cam, sd.render, spp=16, seed=8)[0], ref)
    e4k = relmse(img, ref)
    print(f"  relMSE 16spp {e16:.5f} 4096spp {e4k:.6f} (1/256 of 16spp would be {e16 / 256:.6f})", flush=True)
    # variance falls as 1/spp: 4096 spp should be ~1/256 of 16 spp plus the reference's own noise (1/4096)
    assert e4k < e16 / 100
    assert np.allclose(img.mean(axis=(0, 1)), ref.mean(axis=(0, 1)), rtol=0.01)


def test_deterministic_and_seed_dependent(ctx):
    sd = scenes.cornell(width=32, height=32)
    g = ctx.upload(sd)
    a, _ = gpu_image(g, sd.camera, sd.render, spp=16, seed=3)
    b, _ = gpu_image(g, sd.camera, sd.render, spp=16, seed=3)
    c, _ = gpu_image(g, sd.camera, sd.render, spp=16, seed=4)
    # the film accumulates with float atomics: the order of additions may differ between runs
    assert np.allclose(a, b, rtol=1e-5, atol=1e-6)
    assert not np.allclose(a, c, rtol=1e-3)


def _block_z_test(ctx, sd, cam, spp_cpu, n_cpu, spp_gpu, label, block=8):
    """Bias test robust to fireflies: n_cpu independent CPU renders at spp_cpu give per-block means
    and their standard error; the GPU render at spp_gpu (much less noisy) must sit within the CPU
    error bars: max |z| < 5 and mean z^2 < 3 over blocks and channels (expected 1)."""
    g = ctx.upload(sd)
    o = O.OracleScene(sd)
    print(f"[render] {label}: cpu {n_cpu} x {spp_cpu} spp, gpu {spp_gpu} spp", flush=True)
    H, W = cam["height"], cam["width"]

    def blocks(img):
        return img.reshape(H // block, block, W // block, block, 3).mean(axis=(1, 3))

    cb = np.stack([blocks(cpu_image(o, cam, spp_cpu, s)) for s in range(n_cpu)])
    G, st = gpu_image(g, cam, sd.render, spp=spp_gpu, seed=11)
    gb = blocks(G)
    m = cb.mean(0)
    se = cb.std(0, ddof=1) / np.sqrt(n_cpu) + 1e-4 * (np.abs(m) + 1e-3)
    z = (gb - m) / se
    print(f"  max|z| {np.abs(z).max():.2f} mean z^2 {np.mean(z * z):.2f} mean gpu {G.mean():.5f} cpu {m.mean():.5f} "
          f"nan {st.nan_samples}", flush=True)
    assert np.abs(z).max() < 5.0 and np.mean(z * z) < 3.0
    assert st.nan_samples == 0


def test_glossy_cornell_unbiased(ctx):
    sd = scenes.glossy_cornell()
    cam = dict(sd.camera, width=32, height=32)
    _block_z_test(ctx, sd, cam, 128, 16, 16384, "glossy cornell (dielectric spheres, conductor walls)")


def test_specular_cornell_unbiased(ctx):
    sd = scenes.specular_cornell()
    cam = dict(sd.camera, width=32, height=32)
    # caustics through the glass sphere (diffuse -> glass -> glass -> light, all by BSDF sampling)
    # make the per-block CPU variance heavy-tailed: more CPU renders than for the other scenes
    _block_z_test(ctx, sd, cam, 256, 32, 16384, "specular cornell (smooth mirror box, smooth glass sphere)")


def test_displaced_sphere_env_unbiased(ctx):
    sd = scenes.displaced_sphere(level=5, width=32, height=32)
    _block_z_test(ctx, sd, sd.camera, 64, 8, 8192, "displaced sphere level 5 + env map")


def test_large_scene_render_runs(ctx):
    """The full-size stress scene at low spp: finishes, no NaN, plausible energy."""
    sd = scenes.displaced_sphere(width=256, height=256, spp=4)
    g = ctx.upload(sd)
    print("[render] displaced sphere 1.3M tris 256x256 4spp", flush=True)
    img, st = gpu_image(g, sd.camera, sd.render)
    assert st.samples == 256 * 256 * 4
    assert st.nan_samples == 0
    assert 0.05 < img.mean() < 5.0


FILTERS = [rtg.filter_desc("box"), rtg.filter_desc("box", (1.0, 1.0)), rtg.filter_desc("gaussian", (1.5, 1.5)),
           rtg.filter_desc("mitchell", (2.0, 2.0))]


@pytest.mark.parametrize("flt", FILTERS, ids=["box0.5", "box1", "gauss1.5", "mitchell2"])
def test_filtered_film_matches_oracle_weights(ctx, flt):
    """Screen.cpp:44-158 on a constant scene (uniform environment, no geometry), 40x24 so the
    right and bottom 32x32 blocks are ragged: the resolved image is the constant radiance in every
    pixel, the box of radius 0.5 gives every pixel exactly spp, and the weight film agrees with the
    CPU renderer's (different sample streams: per-pixel weights within 5% on average, totals 1%)."""
    import test_oracle as T
    sd = T._constant_scene(spp=64, flt=flt)
    g = ctx.upload(sd)
    print(f"[render] constant scene 40x24 64spp filter {flt}", flush=True)
    out, st = g.render(sd.camera, sd.render)
    out = np.asarray(out, np.float64)
    w = out[..., 3]
    rgb_c, w_c, _, _ = O.OracleScene(sd).render(threads=min(16, os.cpu_count()))
    img_c = rgb_c / w_c[..., None]
    assert st.nan_samples == 0 and np.all(w > 0)
    img = out[..., :3] / w[..., None]
    assert np.allclose(img, img_c[0, 0][None, None, :], rtol=2e-5), np.abs(img / img_c[0, 0] - 1).max()
    if flt["radius"] == (0.5, 0.5):
        assert np.all(w == 64.0)
    rel = np.abs(w - w_c) / np.abs(w_c)
    print(f"  weight total gpu {w.sum():.1f} cpu {w_c.sum():.1f}; per-pixel rel diff mean {rel.mean():.4f} "
          f"max {rel.max():.4f}", flush=True)
    assert abs(w.sum() / w_c.sum() - 1) < 0.01 and rel.mean() < 0.05


def test_filtered_cornell_unbiased(ctx):
    """Mitchell-Netravali filter (negative lobes) on the Cornell box: block means agree with the CPU
    renderer's using the same filter."""
    sd = scenes.cornell()
    sd.render = dict(sd.render, filter=rtg.filter_desc("mitchell", (2.0, 2.0)))
    cam = dict(sd.camera, width=32, height=32)
    _block_z_test(ctx, sd, cam, 64, 8, 4096, "cornell, Mitchell-Netravali filter radius 2")


def test_gaussian_filter_blurs_an_edge(ctx):
    """A wider filter spreads radiance across pixels: against the box of radius 0.5, the Gaussian of
    radius 1.5 lowers the image's total variation on the Cornell box by more than 8% (edges are smoothed) while the
    mean radiance stays within 2%."""
    sd = scenes.cornell(width=32, height=32)
    g = ctx.upload(sd)
    a, _ = gpu_image(g, sd.camera, sd.render, spp=1024, seed=5)
    b, _ = gpu_image(g, sd.camera, dict(sd.render, filter=rtg.filter_desc("gaussian", (1.5, 1.5))), spp=1024, seed=5)

    def tv(x):
        return np.abs(np.diff(x, axis=0)).sum() + np.abs(np.diff(x, axis=1)).sum()
    print(f"[render] TV box {tv(a):.2f} gaussian {tv(b):.2f}, mean {a.mean():.4f} {b.mean():.4f}", flush=True)
    assert tv(b) < 0.92 * tv(a)   # 0.85 measured at 1024 spp (alpha 2 keeps the Gaussian narrow)
    assert abs(b.mean() / a.mean() - 1) < 0.02


@pytest.mark.parametrize("mode", [rtg.TONEMAP_CLAMP, rtg.TONEMAP_GAMMA, rtg.TONEMAP_REINHARD])
def test_tone_mapping_matches_oracle(ctx, mode, tmp_path):
    """rtg_tonemap on a rendered Cornell film against the oracle's tone mappers: identical bytes
    except where powf rounds differently by an ulp at a truncation boundary (at most 1 code value,
    on under 0.5% of the channels); the PNG written from it reads back unchanged."""
    from raytracer_amd import image
    sd = scenes.cornell(width=64, height=64)
    film, _ = ctx.upload(sd).render(sd.camera, dict(sd.render, spp=64, seed=5))
    for exposure in ((0.5, 0.6) if mode == rtg.TONEMAP_GAMMA else (0.5,)):
        g = ctx.tonemap(film, mode, exposure).astype(int)
        c = O.tonemap(film, mode, exposure).astype(int)
        d = np.abs(g - c)
        print(f"[tonemap] mode {mode} exposure {exposure}: max diff {d.max()}, differing {np.mean(d > 0):.5f}",
              flush=True)
        assert d.max() <= 1 and np.mean(d > 0) < 0.005
    p = str(tmp_path / "img.png")
    image.write_png(p, g.astype(np.uint8))
    assert np.array_equal(image.read_png(p), g.astype(np.uint8))


def test_all_lights_strategy_unbiased(ctx):
    """The all_lights strategy (PathTracingMIS.cpp:119-137: every light sampled once per vertex, one
    shadow ray each) on a Cornell box with three lights (ceiling panel, wall panel, sky): block
    means of the GPU render at 8192 spp within the error bars of 8 CPU renders of 128 spp."""
    sd = scenes.two_light_cornell(width=32, height=32)
    _block_z_test(ctx, sd, sd.camera, 128, 8, 8192, "cornell with three lights, all_lights")


def test_all_lights_matches_one_light(ctx):
    """With a single light, all_lights and one_light_uniform are the same estimator drawing the same
    numbers (selection probability 1, the light's sample uses the same dimension), so the images
    agree up to the order of the film's float additions: this checks the all-lights shadow queue
    (several entries per slot, atomic radiance adds) against the one-light path. (With several
    lights they differ by design: the one-light strategies weight BSDF-sampled emission without the
    selection probability, PathTracingMIS.cpp:205.)"""
    imgs = {}
    for strat in (rtg.LIGHTS_ALL, rtg.LIGHTS_UNIFORM):
        sd = scenes.cornell(width=32, height=32)
        sd.light_strategy = strat
        imgs[strat], st = gpu_image(ctx.upload(sd), sd.camera, sd.render, spp=256, seed=21)
        assert st.nan_samples == 0
    a, b = imgs[rtg.LIGHTS_ALL], imgs[rtg.LIGHTS_UNIFORM]
    print(f"[render] all_lights vs one_light_uniform, one light: max |diff| {np.abs(a - b).max():.3e}", flush=True)
    assert a.mean() > 0 and np.allclose(a, b, rtol=1e-5, atol=1e-6)


def test_zero_light_strategy_is_the_reference_default(ctx):
    """A scene description whose light_strategy is left at 0 (RTG_LIGHTS_DEFAULT, e.g. memset) uses
    the reference's default one_light_uniform (integrator/Integrator.cpp:14): on the three-light
    Cornell box the image equals the one_light_uniform image (same numbers drawn; only the film's
    float addition order may differ) and not the all_lights one."""
    imgs = {}
    for strat in (rtg.LIGHTS_DEFAULT, rtg.LIGHTS_UNIFORM, rtg.LIGHTS_ALL):
        sd = scenes.two_light_cornell(width=24, height=24, strategy=strat)
        imgs[strat], _ = gpu_image(ctx.upload(sd), sd.camera, sd.render, spp=64, seed=22)
    assert np.allclose(imgs[rtg.LIGHTS_DEFAULT], imgs[rtg.LIGHTS_UNIFORM], rtol=1e-5, atol=1e-6)
    assert not np.allclose(imgs[rtg.LIGHTS_DEFAULT], imgs[rtg.LIGHTS_ALL], rtol=1e-3, atol=1e-4)


def test_tuning_and_graphs_leave_the_image_unchanged(ctx):
    """rtg_autotune picks a configuration within its ranges; renders under it, with hipGraph replay
    off, and under a hand-set configuration equal the default render up to the film's float
    addition order (a path's random numbers do not depend on slots, launches or iterations)."""
    sd = scenes.glossy_cornell(width=48, height=32)
    g = ctx.upload(sd)
    base = ctx.tuning()
    ref, _ = gpu_image(g, sd.camera, sd.render, spp=64, seed=23)
    try:
        best, rate = ctx.autotune(g, sd.camera, sd.render, trial_samples=1 << 20)
        print(f"[render] autotune: {best} at {rate / 1e6:.0f} M samples/s", flush=True)
        assert 1 <= best["trace_waves"] <= 8 and 1 <= best["trace_refill"] <= 64 and best["pool_slots"] >= 256
        assert ctx.tuning() == best and rate > 0
        for t in (best, dict(best, graphs=0), dict(base, pool_slots=1 << 16, trace_waves=3, trace_refill=4)):
            ctx.set_tuning(**t)
            img, st = gpu_image(g, sd.camera, sd.render, spp=64, seed=23)
            assert st.samples == 48 * 32 * 64
            assert np.allclose(img, ref, rtol=1e-4, atol=1e-6), t
    finally:
        ctx.set_tuning(**base)


@pytest.mark.parametrize("which", ["cornell", "vol_cornell"])
def test_gather_film_equals_atomic_film(ctx, which):
    """A render whose samples all fit the path pool sums every pixel from the slots after the last
    iteration (k_gather); the same render on a pool smaller than its sample count splats each
    finished sample with float atomics. The films agree up to the order of the float additions and
    the weights (sample counts) exactly."""
    sd = getattr(scenes, which)(width=64, height=48, spp=16)
    g = ctx.upload(sd)
    base = ctx.tuning()
    print(f"[render] {which} 64x48 16spp: gather (2^20 slots) vs atomics (2^14 slots)", flush=True)
    try:
        ctx.set_tuning(**dict(base, pool_slots=1 << 20))
        a, sa = g.render(sd.camera, dict(sd.render, seed=5))
        ctx.set_tuning(**dict(base, pool_slots=1 << 14))
        b, sb = g.render(sd.camera, dict(sd.render, seed=5))
    finally:
        ctx.set_tuning(**base)
    assert sa.samples == sb.samples == 64 * 48 * 16 and sa.nan_samples == sb.nan_samples == 0
    assert sb.iterations > sa.iterations
    assert np.array_equal(a[..., 3], b[..., 3])
    assert np.allclose(a[..., :3], b[..., :3], rtol=1e-5, atol=1e-6)
