"""The renderer's sampler on the GPU (rtg_sampler_batch), against the invariants of the reference's
samplers (sampler/StratifiedSampler.cpp:18-124, RandomSampler.cpp:28-45, Appendix B of the survey's
dimension schedule): with the stratified sampler and N = s^2 samples per pixel, each of the first
`dimension` 2D draws puts exactly one of a pixel's samples in every cell of the s x s grid and each
of the first `dimension` 1D draws one in every interval of [0,1) / N; a non-square N rounds up to
the next square (StratifiedSampler.cpp:23-35); draws beyond `dimension`, and every draw of the
random sampler, are uniform on [0,1) and independent across dimensions; every value lies in [0, 1)
(at most ONE_MINUS_EPS); different pixels use different stratum assignments."""
import numpy as np
import pytest

from raytracer_amd import rtg, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


def _draws(ctx, spp, sampler=rtg.SAMPLER_STRATIFIED, dimension=4, npix=64, n2=7, n1=7, seed=3, width=32):
    sd = scenes.cornell(width=width, height=width)
    rp = dict(sd.render, spp=spp, sampler=sampler, dimension=dimension, seed=seed)
    s = int(np.ceil(np.sqrt(spp)))
    n = s * s if sampler == rtg.SAMPLER_STRATIFIED else spp
    pix = np.arange(npix) * 13 % (width * width)
    ps = np.stack([np.repeat(pix, n), np.tile(np.arange(n), npix)], 1)
    d2, d1 = ctx.sampler_draws(sd.camera, rp, ps, n2, n1)
    # float64 copies: the products with s and n below are then exact (a float32 product rounds)
    return (d2.astype(np.float64).reshape(npix, n, n2, 2), d1.astype(np.float64).reshape(npix, n, n1), s, n)


def _ks_uniform(x):
    """Kolmogorov-Smirnov distance of samples from U[0,1) and its 0.1% critical value."""
    x = np.sort(np.asarray(x, np.float64).ravel())
    m = len(x)
    dplus = np.max(np.arange(1, m + 1) / m - x)
    dminus = np.max(x - np.arange(m) / m)
    return max(dplus, dminus), 1.95 / np.sqrt(m)


@pytest.mark.parametrize("spp", [16, 64, 1])
def test_stratified_dimensions_cover_every_stratum_once(ctx, spp):
    d2, d1, s, n = _draws(ctx, spp)
    assert np.all((d2 >= 0) & (d2 < 1)) and np.all((d1 >= 0) & (d1 < 1))
    for k in range(4):
        cells = np.floor(d2[:, :, k, 0] * s).astype(int) * s + np.floor(d2[:, :, k, 1] * s).astype(int)
        for p in range(cells.shape[0]):
            assert np.array_equal(np.sort(cells[p]), np.arange(n)), (k, p)
        ints = np.floor(d1[:, :, k] * n).astype(int)
        for p in range(ints.shape[0]):
            assert np.array_equal(np.sort(ints[p]), np.arange(n)), (k, p)
    if spp > 1:
        # jitter within the cells is uniform, and pixels differ in their stratum assignment
        off = (d2[:, :, :4, :] * s) % 1.0
        ks, crit = _ks_uniform(off)
        assert ks < crit, (ks, crit)
        c0 = np.floor(d2[:, :, 0, 0] * s) * s + np.floor(d2[:, :, 0, 1] * s)
        assert len({tuple(r) for r in c0.astype(int)}) > c0.shape[0] // 2


@pytest.mark.parametrize("spp,rounded", [(10, 16), (20, 25), (4099, 4225)])
def test_non_square_count_rounds_up_and_stays_stratified(ctx, spp, rounded):
    d2, d1, s, n = _draws(ctx, spp, npix=4)
    assert n == rounded and s * s == rounded
    for k in range(4):
        cells = np.floor(d2[:, :, k, 0] * s).astype(int) * s + np.floor(d2[:, :, k, 1] * s).astype(int)
        ints = np.floor(d1[:, :, k] * n).astype(int)
        for p in range(cells.shape[0]):
            assert np.array_equal(np.sort(cells[p]), np.arange(n))
            assert np.array_equal(np.sort(ints[p]), np.arange(n))


def test_draws_beyond_the_stratified_dimensions_are_uniform_and_independent(ctx):
    d2, d1, s, n = _draws(ctx, 64, npix=128, n2=8, n1=8)
    beyond2 = d2[:, :, 4:, :]
    beyond1 = d1[:, :, 4:]
    for x in (beyond2, beyond1):
        ks, crit = _ks_uniform(x)
        assert ks < crit, (ks, crit)
    # no stratification beyond dimension: some pixel has two samples in one cell
    cells = np.floor(beyond2[:, :, 0, 0] * s) * s + np.floor(beyond2[:, :, 0, 1] * s)
    assert any(len(np.unique(r)) < n for r in cells)
    # independence across dimensions (a stratified and a raw dimension, two raw ones)
    flat = lambda a: a.reshape(-1)
    for a, b in ((d2[:, :, 0, 0], d2[:, :, 5, 0]), (d2[:, :, 5, 0], d2[:, :, 6, 1]), (d1[:, :, 1], d1[:, :, 6])):
        r = np.corrcoef(flat(a), flat(b))[0, 1]
        assert abs(r) < 4 / np.sqrt(flat(a).size), r


def test_dimension_parameter_sets_the_stratified_count(ctx):
    """dimension = 2: draws 0-1 stratified, draw 2 raw."""
    d2, d1, s, n = _draws(ctx, 16, dimension=2, npix=256, n2=3, n1=3)
    for k in range(2):
        cells = np.floor(d2[:, :, k, 0] * s) * s + np.floor(d2[:, :, k, 1] * s)
        assert all(len(np.unique(r)) == n for r in cells)
    cells = np.floor(d2[:, :, 2, 0] * s) * s + np.floor(d2[:, :, 2, 1] * s)
    assert sum(len(np.unique(r)) < n for r in cells) > 200


def test_random_sampler_is_uniform(ctx):
    d2, d1, s, n = _draws(ctx, 64, sampler=rtg.SAMPLER_RANDOM, npix=128, n2=4, n1=4)
    for x in (d2, d1):
        assert np.all((x >= 0) & (x < 1))
        ks, crit = _ks_uniform(x)
        assert ks < crit
    cells = np.floor(d2[:, :, 0, 0] * 8) * 8 + np.floor(d2[:, :, 0, 1] * 8)
    assert sum(len(np.unique(r)) < n for r in cells) > 100  # not stratified
