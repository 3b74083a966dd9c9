"""Image textures on the GPU (texture/ImageTexture.cpp, ScaleTexture.cpp) against the CPU oracle:
a closed form through whole renders, block means of the textured Cornell box against CPU renders,
and the scene-file round trip."""
import os

import numpy as np
import pytest

from raytracer_amd import rtg, scenes
import oracle_lib as O
import test_oracle_texture as TT
from test_gpu_render import _block_z_test, gpu_image
from test_xml_host import decode_rgbe

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


def test_textured_plane_under_a_uniform_sky(ctx):
    """The oracle test's closed form on the GPU: with the same seed the paths do not depend on the
    albedo, so the textured floor's image is the texture at each pixel centre's uv (from the oracle)
    times the white floor's image; mean absolute error below 1% of the mean at 256 spp."""
    W, H = 16, 12
    rgb = np.ascontiguousarray(TT._linear(W, H, a=0.1, b=0.04, c=0.03))
    v, i = scenes.quad((-2, 0, -2), (4, 0, 0), (0, 0, 4), (0, 1, 0))
    uvq = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    floor = scenes.MeshDesc(v, i, material=0, uvs=uvq)
    env = np.ones((16, 32, 3), np.float32)
    sd = TT._tex_scene([dict(rgb=rgb)], [dict(scenes.diffuse((0.5,) * 3), tex_albedo=1)], [floor], env=env)
    white = TT._tex_scene([dict(rgb=rgb)], [scenes.diffuse((1.0, 1.0, 1.0))], [floor], env=env)
    n = 32
    cam = dict(origin=(0.0, 2.0, 0.0), target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), fov=40.0, width=n, height=n)
    rp = dict(spp=256, sampler=rtg.SAMPLER_STRATIFIED, dimension=4, min_depth=3, max_depth=4, seed=2)
    print("[texture] textured floor under a uniform sky, 32x32 256 spp", flush=True)
    img, st = gpu_image(ctx.upload(sd), cam, rp)
    base, _ = gpu_image(ctx.upload(white), cam, rp)
    o = O.OracleScene(sd)
    rays, _ = O.camera_rays(cam, np.array([[x + 0.5, y + 0.5] for y in range(n) for x in range(n)], np.float32))
    _, uv = o.material_at(rays)
    want = o.texture_eval(0, uv).reshape(n, n, 3) * base
    err = np.abs(img - want).mean() / want.mean()
    print(f"  mean |err| / mean {err:.5f}", flush=True)
    assert st.nan_samples == 0 and err < 0.01


def test_textured_cornell_unbiased(ctx):
    """Checkerboard back wall (mesh uvs), smooth-ramp short box (barycentric uvs, uv scale 2, value
    scale 0.9), rough-conductor sphere with a roughness texture (spherical uvs, inverted y): block
    means of the GPU render at 8192 spp within the error bars of 16 CPU renders of 256 spp (the
    glossy sphere's reflection of the light makes the per-block CPU estimates heavy-tailed)."""
    sd = scenes.textured_cornell(width=32, height=32)
    _block_z_test(ctx, sd, sd.camera, 256, 16, 8192, "textured cornell")


def test_textured_scene_file_round_trip(ctx, tmp_path):
    """The textured Cornell box written as a scene file (textures as RGBE .hdr images, uvs in the OBJ
    files) renders, with the same seed, like the direct upload of the same scene with the textures
    quantised to RGBE: equal up to the film's atomic order."""
    sd = scenes.textured_cornell(width=32, height=32, spp=64)
    path = scenes.write_xml(sd, str(tmp_path))
    xs, xcam, xrp = ctx.load_xml(path)
    for k, t in enumerate(sd.textures):
        t["rgb"] = decode_rgbe(os.path.join(str(tmp_path), f"tex{k + 1}.hdr")).astype(np.float32)
    print("[texture] textured cornell 32x32 64 spp, file vs direct upload", flush=True)
    a, _ = gpu_image(xs, xcam, dict(xrp, spp=64, seed=7))
    b, _ = gpu_image(ctx.upload(sd), sd.camera, dict(sd.render, spp=64, seed=7))
    d = np.abs(a - b)
    print(f"  max |diff| {d.max():.3e}, mean {a.mean():.5f}", flush=True)
    assert a.mean() > 0 and np.all(d <= 1e-5 * np.maximum(1.0, b)), d.max()
