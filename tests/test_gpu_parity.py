"""GPU parity against the CPU oracle: first hits, BSDF eval/pdf/sample, light sampling, images."""
import numpy as np
import pytest

from raytracer_amd import rtg, scenes
import oracle_lib as O
from gpu_util import random_rays, compare_hits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


def _trace_parity(ctx, sd, rays, label):
    print(f"[trace parity] {label}: {rays.shape[0]} rays", flush=True)
    g = ctx.upload(sd)
    o = O.OracleScene(sd)
    pg, tg, uvg, ppg, ng = g.trace(rays)
    pb, tb, uvb, ppb, nb = o.trace(rays, brute=True)
    po, to, uvo, ppo, no = o.trace(rays)
    same, tie, bad = compare_hits(pg, tg, pb, tb)
    assert bad.sum() == 0, f"{label}: {bad.sum()} non-tie mismatches vs brute force"
    assert tie.sum() <= max(1, int(1e-5 * len(rays))) + 2, f"{label}: {tie.sum()} ties"
    m = same & (pg >= 0)
    # same primitive -> bit-identical t, barycentrics, point and shading normal
    assert np.array_equal(tg[m].view(np.uint32), tb[m].view(np.uint32))
    assert np.array_equal(uvg[m].view(np.uint32), uvb[m].view(np.uint32))
    assert np.array_equal(ppg[m].view(np.uint32), ppb[m].view(np.uint32))
    assert np.array_equal(ng[m].view(np.uint32), nb[m].view(np.uint32))
    # the reference-structured BVH agrees with brute force too
    s2, t2, b2 = compare_hits(po, to, pb, tb)
    assert b2.sum() == 0
    # any-hit: occlusion flags equal
    ag = g.trace(rays, any_hit=True)[0] >= 0
    ao = o.trace(rays, any_hit=True, brute=True)[0] >= 0
    assert np.array_equal(ag, ao), f"{label}: any-hit mismatch {np.sum(ag != ao)}"
    g.release()
    return same.mean(), tie.sum()


def test_trace_cornell(ctx):
    sd = scenes.cornell()
    rng = np.random.default_rng(1)
    rays = random_rays(rng, 200000, np.array([0, 0.5, 0]), 0.45)
    cam = sd.camera
    pxy = rng.random((50000, 2)) * [cam["width"], cam["height"]]
    crays, _ = O.camera_rays(cam, pxy)
    _trace_parity(ctx, sd, np.concatenate([rays, crays]), "cornell")


def test_trace_spheres(ctx):
    sd = scenes.glossy_cornell()
    rng = np.random.default_rng(2)
    rays = random_rays(rng, 100000, np.array([0, 0.5, 0]), 0.45)
    _trace_parity(ctx, sd, rays, "glossy cornell (spheres)")


def test_trace_displaced_small(ctx):
    sd = scenes.displaced_sphere(level=5, width=64, height=64)
    rng = np.random.default_rng(3)
    rays = random_rays(rng, 100000, np.array([0, 1.1, 0]), 1.3)
    _trace_parity(ctx, sd, rays, "displaced sphere level 5")


def test_trace_displaced_full_size(ctx):
    """The 1,310,722-triangle benchmark scene (level-8 displaced icosphere + ground): 100k camera
    rays of the 1024x1024 benchmark camera, 100k random rays and 100k secondary rays leaving the
    camera rays' hit points in random directions (adaptive epsilon). GPU first hits against the
    oracle's reference-structured BVH (itself checked against brute force on a 1,500-ray subset):
    equal ids except near-ties (at most 1e-5 of rays), bit-identical t, barycentrics, point and
    normal; any-hit flags equal. This tree is deep enough that traversal stacks spill past their
    on-chip (LDS) entries: the spill count of the trace is asserted to be non-zero, so the global
    spill path is covered by the parity check."""
    sd = scenes.displaced_sphere(level=8)
    rng = np.random.default_rng(4)
    cam = sd.camera
    crays, _ = O.camera_rays(cam, rng.random((100000, 2)) * [cam["width"], cam["height"]])
    rrays = random_rays(rng, 100000, np.array([0, 1.1, 0]), 1.3)
    o = O.OracleScene(sd)
    pc, tc, _, ptc, _ = o.trace(crays)
    hitp = ptc[pc >= 0][:100000]
    srays = random_rays(rng, hitp.shape[0], np.zeros(3), 0.0)
    srays[:, :3] = hitp
    rays = np.concatenate([crays, rrays, srays])
    print(f"[trace parity] displaced sphere level 8 ({sd.num_triangles()} triangles): {len(rays)} rays", flush=True)
    g = ctx.upload(sd)
    pg, tg, uvg, ppg, ng = g.trace(rays)
    spills = ctx.trace_spills()
    po, to, uvo, ppo, no = o.trace(rays)
    same, tie, bad = compare_hits(pg, tg, po, to)
    print(f"  equal ids {same.mean():.6f}, ties {tie.sum()}, stack spills {spills}", flush=True)
    assert bad.sum() == 0, f"{bad.sum()} non-tie mismatches"
    assert tie.sum() <= max(1, int(1e-5 * len(rays))), f"{tie.sum()} ties"
    assert spills > 0
    m = same & (pg >= 0)
    assert m.sum() > len(rays) // 3
    assert np.array_equal(tg[m].view(np.uint32), to[m].view(np.uint32))
    assert np.array_equal(uvg[m].view(np.uint32), uvo[m].view(np.uint32))
    assert np.array_equal(ppg[m].view(np.uint32), ppo[m].view(np.uint32))
    assert np.array_equal(ng[m].view(np.uint32), no[m].view(np.uint32))
    sub = rng.choice(len(rays), 1500, replace=False)
    pb, tb = o.trace(rays[sub], brute=True)[:2]
    assert compare_hits(po[sub], to[sub], pb, tb)[2].sum() == 0
    ag = g.trace(rays, any_hit=True)[0] >= 0
    ao = o.trace(rays, any_hit=True)[0] >= 0
    assert np.array_equal(ag, ao), f"any-hit mismatch {np.sum(ag != ao)}"
    g.release()


def _dirs(rng, n, upper=None):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if upper is True:
        d[:, 2] = np.abs(d[:, 2])
    return d.astype(np.float32)


MATS = [
    scenes.diffuse((0.73, 0.5, 0.2)),
    scenes.conductor(0.15),
    scenes.conductor(0.4, dist=rtg.DIST_BECKMANN),
    scenes.dielectric(0.1),
    scenes.dielectric(0.3, dist=rtg.DIST_BECKMANN),
    scenes.smooth_conductor(eta=(0.2, 0.92, 1.1), k=(3.9, 2.45, 2.14), reflectance=(0.9, 0.8, 0.7)),
    scenes.smooth_dielectric(eta_int=1.5, eta_ext=1.000277, transmitted=(0.9, 0.95, 1.0)),
]


@pytest.mark.parametrize("mi", range(len(MATS)))
def test_bsdf_eval_pdf(ctx, mi):
    m = MATS[mi]
    print(f"[bsdf eval] material {mi}", flush=True)
    sd = scenes.SceneDesc(materials=[m])
    o = O.OracleScene(sd)
    rng = np.random.default_rng(10 + mi)
    n = 100000
    wi = _dirs(rng, n); wo = _dirs(rng, n)
    fg, pg = ctx.bsdf_eval(m, wi, wo)
    fo, po = o.bsdf_eval(0, wi, wo)
    # GGX eval/pdf use only + - * / sqrt (bit-exact); Beckmann's D adds exp(), a few ulp apart
    # between the device and glibc, and inexact near the bottom of the float range where the
    # device's exp goes through subnormal intermediates. Stated tolerance: 1e-5 relative wherever
    # the CPU value is >= 1e-30.
    tiny = 1e-30
    nf = np.abs(fo) >= tiny
    npd = np.abs(po) >= tiny
    rel_f = np.abs(fg - fo)[nf] / np.abs(fo)[nf]
    rel_p = np.abs(pg - po)[npd] / np.abs(po)[npd]
    print(f"  max rel f {rel_f.max() if rel_f.size else 0:.3e} pdf {rel_p.max() if rel_p.size else 0:.3e}; "
          f"bit-exact f={np.mean(fg.view(np.uint32) == fo.view(np.uint32)):.6f} "
          f"pdf={np.mean(pg.view(np.uint32) == po.view(np.uint32)):.6f}; subnormal refs {np.sum(~nf)}")
    assert np.all(rel_f <= 1e-5) and np.all(rel_p <= 1e-5)
    assert np.all(np.abs(fg[~nf]) < 1e-30) and np.all(np.abs(pg[~npd]) < 1e-30)
    assert np.array_equal(np.isnan(fg), np.isnan(fo)) and np.array_equal(np.isnan(pg), np.isnan(po))
    if m.get("distribution", 0) == rtg.DIST_GGX or m["type"] == rtg.MAT_DIFFUSE:
        assert np.array_equal(fg.view(np.uint32), fo.view(np.uint32))
        assert np.array_equal(pg.view(np.uint32), po.view(np.uint32))


@pytest.mark.parametrize("mi", range(len(MATS)))
def test_bsdf_sample(ctx, mi):
    m = MATS[mi]
    print(f"[bsdf sample] material {mi}", flush=True)
    sd = scenes.SceneDesc(materials=[m])
    o = O.OracleScene(sd)
    rng = np.random.default_rng(20 + mi)
    n = 100000
    wi = _dirs(rng, n, upper=(m["type"] not in (rtg.MAT_ROUGH_DIELECTRIC, rtg.MAT_DIELECTRIC)))
    u = rng.random((n, 2)).astype(np.float32)
    wg, fg, pg, eg = ctx.bsdf_sample(m, wi, u)
    wo, fo, po, eo, _ = o.bsdf_sample(0, wi, u)
    okg = np.any(fg != 0, axis=1); oko = np.any(fo != 0, axis=1)
    # sampling uses sin/cos/atan/log, which differ between the device and glibc by ulps: the
    # outcome (zero / non-zero) differs only where a decision sits on a rounding boundary
    agree = okg == oko
    assert agree.mean() > 0.9999, agree.mean()
    b = okg & oko
    assert np.allclose(wg[b], wo[b], rtol=0, atol=2e-5)
    rel = np.abs(fg[b] - fo[b]) / np.maximum(np.abs(fo[b]), 1e-6)
    assert np.quantile(rel, 0.999) < 1e-3
    relp = np.abs(pg[b] - po[b]) / np.maximum(np.abs(po[b]), 1e-6)
    assert np.quantile(relp, 0.999) < 1e-3
