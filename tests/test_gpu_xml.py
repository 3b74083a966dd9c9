"""Scenes loaded from files in the reference's XML schema (rtg_scene_load_xml) render like the same
scenes uploaded directly.

scenes.write_xml writes every float with its exact decimal form and keeps triangles in the same order
where each mesh's materials are contiguous (Cornell box, glossy Cornell), so those scenes load back
bit-for-bit equal and, with the same seed, trace the same paths: the camera frame of the file's
<lookAt> is built with the same double-precision steps as the library's own. The images then agree
up to the order in which the film's float atomic adds land (a few ulp; any path that differs would
move a pixel by orders of magnitude more). The displaced sphere interleaves materials over its
triangles (so the file reorders them, and the BVH differs) and its environment map goes through
8-bit RGBE: it is compared with the direct upload of the same quantised map, where only exact
first-hit ties may take another path."""
import os

import numpy as np
import pytest

from raytracer_amd import rtg, scenes
from test_xml_host import decode_rgbe

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


def _render(scene, cam, rp):
    out, st = scene.render(cam, rp)
    assert st.nan_samples == 0
    return rtg.resolve(out).astype(np.float64)


def _pair(ctx, sd, tmp_path, spp):
    path = scenes.write_xml(sd, str(tmp_path))
    xs, xcam, xrp = ctx.load_xml(path)
    ps = ctx.upload(sd)
    rp = dict(sd.render, spp=spp, seed=7)
    assert (xrp["min_depth"], xrp["max_depth"], xrp["sampler"]) == (rp["min_depth"], rp["max_depth"], rp["sampler"])
    a = _render(xs, xcam, dict(xrp, spp=spp, seed=7))
    b = _render(ps, sd.camera, rp)
    return a, b


@pytest.mark.parametrize("name", ["cornell", "glossy_cornell"])
def test_xml_scene_renders_the_same_paths(ctx, tmp_path, name):
    sd = getattr(scenes, name)(width=64, height=64, spp=256)
    print(f"[xml] {name} 64x64 256 spp, file vs direct upload", flush=True)
    a, b = _pair(ctx, sd, tmp_path, 256)
    d = np.abs(a - b)
    print(f"  max |diff| {d.max():.3e}, differing pixels {np.mean(d.max(axis=2) > 0):.4f}, mean {a.mean():.5f}",
          flush=True)
    assert a.mean() > 0
    assert np.all(d <= 1e-5 * np.maximum(1.0, b)), d.max()


def test_xml_env_scene_matches(ctx, tmp_path):
    sd = scenes.displaced_sphere(width=64, height=64, spp=64, level=4)
    path = scenes.write_xml(sd, str(tmp_path / "x"))
    sd.env = decode_rgbe(os.path.join(os.path.dirname(path), "env.hdr")).astype(np.float32)
    print("[xml] displaced sphere (level 4) 64x64 64 spp with RGBE env, file vs direct upload", flush=True)
    xs, xcam, xrp = ctx.load_xml(path)
    info = xs.info()
    assert (info["env_height"], info["env_width"]) == sd.env.shape[:2]
    a = _render(xs, xcam, dict(xrp, spp=64, seed=3))
    b = _render(ctx.upload(sd), sd.camera, dict(sd.render, spp=64, seed=3))
    d = np.abs(a - b).max(axis=2)
    frac = float(np.mean(d > 1e-5 * np.maximum(1.0, b.max(axis=2))))
    rel = abs(a.mean() - b.mean()) / b.mean()
    print(f"  differing pixels {frac:.4f}, relative mean difference {rel:.2e}", flush=True)
    assert frac < 0.02 and rel < 2e-3


def test_xml_volumetric_scene_renders_the_same_paths(ctx, tmp_path):
    """The volumetric Cornell box (64^3 fBm grid) written with its GridVolume file and read back:
    the loaded grid gives the same densities as the uploaded one (the writer pads the value the
    reference's reader skips) and, with the same seed, the same image up to the film's atomic order."""
    sd = scenes.vol_cornell(width=32, height=32, spp=64, res=64)
    print("[xml] volumetric cornell 32x32 64 spp, file vs direct upload", flush=True)
    a, b = _pair(ctx, sd, tmp_path, 64)
    xs, _, xrp = ctx.load_xml(os.path.join(str(tmp_path), "scene.xml"))
    assert xrp["integrator"] == rtg.INTEGRATOR_VOLPATH
    pts = np.random.default_rng(2).uniform(-0.05, 1.05, (4096, 3)).astype(np.float32)
    assert np.array_equal(xs.medium(0, pts), ctx.upload(sd).medium(0, pts))
    d = np.abs(a - b)
    print(f"  max |diff| {d.max():.3e}, mean {a.mean():.5f}", flush=True)
    assert a.mean() > 0
    assert np.all(d <= 1e-5 * np.maximum(1.0, b)), d.max()


def test_xml_named_metal_matches_explicit_colors(ctx, tmp_path):
    """The mirror of the specular Cornell box named 'gold' (IOR spectra in data/IOR beside the
    scene file) renders, with the same seed, like the same scene with its eta and absorption given
    as the RGB that rtg_spectrum_to_rgb makes of those spectra (the header line of the k file read
    as (0, 0))."""
    import re

    from test_xml_host import named_metal_files

    tmp = str(tmp_path)
    named_metal_files(tmp)
    eta = rtg.spectrum_to_rgb(np.array([400, 500, 600, 700], np.float32), np.array([1.5, 0.9, 0.3, 0.2], np.float32))
    k = rtg.spectrum_to_rgb(np.array([0, 0, 400, 550, 700], np.float32), np.array([0, 0, 1.9, 2.4, 4.0], np.float32))
    col = lambda n, c: f'<color name="{n}" value="{" ".join("%.9g" % x for x in c)}"/>'
    text = open(scenes.write_xml(scenes.specular_cornell(width=32, height=32), tmp)).read()
    pat = re.compile(r'<color name="eta" value="[^"]*"/>\s*<color name="absorption" value="[^"]*"/>')
    assert len(pat.findall(text)) == 1
    paths = []
    for name, repl in (("named.xml", '<string name="material" value="gold"/>'),
                       ("explicit.xml", col("eta", eta) + col("absorption", k))):
        paths.append(os.path.join(tmp, name))
        with open(paths[-1], "w") as f:
            f.write(pat.sub(repl, text))
    print(f"[xml] named metal: eta {eta}, k {k}", flush=True)
    a_sc, cam, rp = ctx.load_xml(paths[0])
    b_sc, _, _ = ctx.load_xml(paths[1])
    rp = dict(rp, spp=64, seed=3)
    a, b = _render(a_sc, cam, rp), _render(b_sc, cam, rp)
    print(f"  mean {a.mean():.5f}, max |diff| {np.abs(a - b).max():.3e}", flush=True)
    assert a.mean() > 0 and np.all(np.abs(a - b) <= 1e-5 * np.maximum(1.0, b))


def test_reference_cbox_fog_file_renders_like_the_oracle(ctx, tmp_path):
    """The reference's own cbox_fog.xml, loaded verbatim (stand-in meshes beside it, see
    tests/reference_standins.py): its VolPathTracing render in the homogeneous fog, with its random
    sampler, Mitchell filter and mirrored camera, at 64x64 and 4096 spp, sits within the error bars
    of 8 oracle renders of the scene the file describes (block means, 8x8 blocks)."""
    import oracle_lib as O
    import reference_standins as RS

    sc, cam, rp = ctx.load_xml(RS.setup_cbox_fog(str(tmp_path)))
    cam = dict(cam, width=64, height=64)
    ref = RS.cbox_fog_scene(width=64, height=64, min_depth=rp["min_depth"], max_depth=rp["max_depth"])
    o = O.OracleScene(ref)
    print(f"[xml] cbox_fog.xml (reference file): depth {rp['min_depth']}..{rp['max_depth']}, "
          f"cpu 8 x 64 spp, gpu 4096 spp", flush=True)

    def blocks(img):
        return img.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3))

    cpu = []
    for s in range(8):
        rgb, w, _, _ = o.render(camera=ref.camera, spp=64, threads=min(16, os.cpu_count()), sampler=0, seed_offset=s)
        cpu.append(blocks(rgb / np.maximum(w[..., None], 1e-30)))
    cb = np.stack(cpu)
    out, st = sc.render(cam, dict(rp, spp=4096, seed=17))
    G = rtg.resolve(out).astype(np.float64)
    m = cb.mean(0)
    se = cb.std(0, ddof=1) / np.sqrt(8) + 1e-4 * (np.abs(m) + 1e-3)
    z = (blocks(G) - m) / se
    print(f"  max|z| {np.abs(z).max():.2f} mean z^2 {np.mean(z * z):.2f} mean gpu {G.mean():.5f} cpu {m.mean():.5f}",
          flush=True)
    assert G.mean() > 0 and st.nan_samples == 0
    assert np.abs(z).max() < 5.0 and np.mean(z * z) < 3.0
