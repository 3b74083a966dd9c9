"""Edge cases of the render path on the GPU against the CPU oracle: a scene with nothing in it
(every camera ray misses, no light, no environment), a single-pixel image, and an image whose
sides are not multiples of the 8x8 claim tile rendered at a sample count that is not a square
(the stratified sampler rounds it up, samplers/StratifiedSampler.cpp)."""
import numpy as np
import pytest

from raytracer_amd import rtg, scenes
import oracle_lib as O
from test_gpu_render import _block_z_test, gpu_image

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


def test_empty_scene_is_black(ctx):
    """No primitives, no lights, no environment: every sample is zero on both sides, none NaN."""
    sd = scenes.SceneDesc()
    sd.add_material(**scenes.diffuse((0.5, 0.5, 0.5)))
    cam = dict(origin=(0.0, 0.0, 5.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0, width=33, height=17)
    rp = dict(spp=4, sampler=rtg.SAMPLER_STRATIFIED, dimension=4, min_depth=3, max_depth=8, seed=1)
    sc = ctx.upload(sd)
    assert sc.info()["triangles"] == 0 and sc.info()["lights"] == 0
    print("[edges] empty scene 33x17 4 spp", flush=True)
    out, st = sc.render(cam, rp)
    assert st.nan_samples == 0 and np.all(out[..., :3] == 0)
    assert np.allclose(out[..., 3], out[0, 0, 3]) and out[0, 0, 3] > 0  # every pixel got its samples
    sd.camera, sd.render = cam, rp
    rgb, w, _, _ = O.OracleScene(sd).render(threads=8)
    assert np.all(rgb == 0) and np.allclose(w, out[..., 3], rtol=1e-5)


def test_single_pixel_image(ctx):
    """A 1x1 Cornell box: the GPU pixel at 65536 spp within the error bar of 16 CPU renders."""
    sd = scenes.cornell(width=1, height=1)
    _block_z_test(ctx, sd, sd.camera, 4096, 16, 65536, "cornell 1x1", block=1)


def test_ragged_image_and_non_square_spp(ctx):
    """A 20x12 Cornell box (partial claim tiles on both axes) at 4099 spp on the GPU against 16
    CPU renders at 50 spp (both rounded up to squares by the stratified sampler), 4x4 blocks."""
    sd = scenes.cornell(width=20, height=12)
    _block_z_test(ctx, sd, sd.camera, 50, 16, 4099, "cornell 20x12 ragged", block=4)
