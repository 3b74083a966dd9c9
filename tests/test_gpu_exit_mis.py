"""Rough-dielectric exit proposals of the guided modes (rtg_render_desc.guiding = RTG_GUIDING_EXIT,
and inside RTG_GUIDING_LEARNED; no counterpart in the reference).

At a rough-dielectric vertex whose incoming direction lies on the denser side of a transmitting
macro-surface, the renderer combines the reference's sampler (bsdf/RoughDielectric.cpp:194-258,
MicrofacetDistribution.cpp:33-48) with two proposals for the microfacet normal, one near the
critical angle and one near grazing, by one-sample MIS in the space of normals. Checked here through
the rtg_bsdf_batch tap (op 3), against an independent float64 restatement of the proposals and the
oracle's sampler (oracle/rt_oracle.cpp):
  * each sample's normal, rebuilt in numpy from (technique, ux, uy), fed to the ORACLE's sampler
    (with the ux the sampler needs for that normal) gives the GPU's direction and, times the MIS
    factor D_a(n) n.z / sum_k P_k p_k(n) computed in numpy, the GPU's estimate;
  * the proposals' sampling maps have the densities the factor uses (uniform azimuth by
    Kolmogorov-Smirnov, the band coordinate's bin counts against the densities' integrals by
    chi-square);
  * the estimator's mean equals the plain sampler's mean (the expectation is kept) while its
    largest weights are bounded where the plain sampler's reach 10^3;
  * elsewhere (entering rays, total internal reflection at the macro-surface) the plain sampler
    runs unchanged, bit for bit;
  * renders: EXIT and LEARNED renders of the glossy-indirect scene agree with the committed
    2^20-spp plain reference (tests/golden/glossy_cornell_ref.npz) block by block."""
import os

import numpy as np
import pytest

from raytracer_amd import rtg, scenes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
P1, P2, B1, G1 = 0.05, 0.10, 0.1, 0.1   # the proposals' constants (rtg_device.h)
ETA_INT, ETA_EXT = 1.5, 1.000277


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc():
    import oracle_lib as O
    mats = [scenes.dielectric(0.1), scenes.dielectric(0.3, dist=rtg.DIST_BECKMANN)]
    return O.OracleScene(scenes.SceneDesc(materials=mats)), mats


def make_frame(n):
    """The renderer's orthonormal frame around n (its `make_frame`), in float64."""
    n = np.asarray(n, np.float64)
    if abs(n[0]) > abs(n[1]):
        inv = 1.0 / np.sqrt(n[0] ** 2 + n[2] ** 2)
        t = np.array([n[2] * inv, 0.0, -n[0] * inv])
    else:
        inv = 1.0 / np.sqrt(n[1] ** 2 + n[2] ** 2)
        t = np.array([0.0, n[2] * inv, -n[1] * inv])
    return np.cross(t, n), t


def adj_alpha(alpha, ci):
    return (1.2 - 0.2 * np.sqrt(abs(ci))) * alpha


def mf_sample(dist, ux, uy, a):
    if dist == rtg.DIST_BECKMANN:
        th = np.arctan(np.sqrt(-a * a * np.log1p(-ux)))
    else:
        th = np.arctan(a * np.sqrt(ux) / np.sqrt(1 - ux))
    ph = 2 * np.pi * uy
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)


def mf_density(dist, n, a):
    """D_a(n) cos(theta_n): the density of mf_sample's normals (solid angle)."""
    nz = n[:, 2]
    c2 = np.maximum(nz * nz, 1e-300)
    t2 = (1 - nz * nz) / c2
    if dist == rtg.DIST_BECKMANN:
        D = np.exp(-t2 / (a * a)) / (np.pi * a * a * c2 * c2)
    else:
        D = a * a / (np.pi * c2 * c2 * (a * a + t2) ** 2)
    return np.where(nz > 0, D * nz, 0.0)


def u_of(dist, n, a):
    """The ux for which mf_sample returns n's polar angle."""
    nz = n[:, 2]
    t2 = (1 - nz * nz) / np.maximum(nz * nz, 1e-300)
    u = -np.expm1(-t2 / (a * a)) if dist == rtg.DIST_BECKMANN else t2 / (a * a + t2)
    return np.minimum(u, 1 - 2.0 ** -24)


def band_normals(wi, tech, ux, uy, cos_c):
    c = np.where(tech == 1, -np.minimum(cos_c + B1 * ux * ux, 1.0), G1 * ux * ux)
    sc = np.sqrt(np.maximum(0, 1 - c * c))
    s, t = make_frame(wi)
    ph = 2 * np.pi * uy
    return (sc * np.cos(ph))[:, None] * s + (sc * np.sin(ph))[:, None] * t + c[:, None] * wi


def band_densities(wi, n, cos_c):
    c = n @ wi
    x = -c - cos_c
    x = np.where((x <= 1e-10) & (x > -1e-5), 1e-10, x)   # the edge floors of the device code
    p1 = np.where((x > 0) & (x <= B1), 1 / (2 * np.sqrt(B1 * np.maximum(x, 1e-300))) / (2 * np.pi), 0.0)
    p2 = np.where((c > -1e-7) & (c <= G1), 1 / (2 * np.sqrt(G1 * np.maximum(c, 1e-12))) / (2 * np.pi), 0.0)
    return p1, p2


def _wi(cz):
    return np.array([np.sqrt(1 - cz * cz), 0.0, cz])


@pytest.mark.parametrize("mi,cz", [(0, -0.95), (0, -0.85), (0, -0.78), (1, -0.9)])
def test_exit_proposals_against_oracle_sampler(ctx, orc, mi, cz):
    O, mats = orc
    mat = mats[mi]
    dist = mat["distribution"]
    n_s = 200_000
    rng = np.random.default_rng(11 + mi)
    u = rng.random((n_s, 3)).astype(np.float32)
    wi = _wi(cz)
    wo, f, pdf, k = ctx.exit_mis_sample(mat, np.tile(wi, (n_s, 1)), u)
    u64 = u.astype(np.float64)
    us, ux, uy = u64[:, 0], u64[:, 1], u64[:, 2]
    tech = np.where(us < P1, 1, np.where(us < P1 + P2, 2, 0))
    a = adj_alpha(mat["alpha"], np.float32(cz))
    cos_c = np.sqrt(1 - (ETA_EXT / ETA_INT) ** 2)
    nb = band_normals(wi, tech, ux, uy, cos_c)
    n = np.where((tech == 0)[:, None], mf_sample(dist, ux, uy, a), nb)
    ps = mf_density(dist, n, a)
    p1, p2 = band_densities(wi, n, cos_c)
    den = (1 - P1 - P2) * ps + P1 * p1 + P2 * p2
    factor = np.where(den > 0, ps / np.maximum(den, 1e-300), 0.0)
    # the oracle's sampler for the same normal
    ueq = np.stack([np.where(tech == 0, ux, u_of(dist, n, a)),
                    np.where(tech == 0, uy, (np.arctan2(n[:, 1], n[:, 0]) / (2 * np.pi)) % 1.0)], 1)
    ueq32 = ueq.astype(np.float32)
    wo_o, w_o, pdf_o, _, _ = O.bsdf_sample(mi, np.tile(wi, (n_s, 1)).astype(np.float32), ueq32)
    # a steep proposed normal needs ux within a few float32 steps of 1, so the oracle's normal for
    # the rounded ux can differ from the proposal's: such samples are not comparable (counted)
    n_o = mf_sample(dist, ueq32[:, 0].astype(np.float64), ueq32[:, 1].astype(np.float64), a)
    same_n = np.einsum("ij,ij->i", n_o, n) > np.cos(2e-5)
    print(f"  normals reproduced by the oracle's ux: {same_n.mean():.5f} "
          f"(sampler {same_n[tech == 0].mean():.5f}, bands {same_n[tech > 0].mean():.5f})")
    assert same_n[tech == 0].mean() > 0.9999
    # the proposals' normals the oracle cannot reach are the steepest, where D_a (hence the MIS
    # factor) is tiny; their estimates are covered by the statistical test below
    assert np.all(factor[~same_n & (n[:, 2] > 0)] < 0.02)
    live = (n[:, 2] > 0) & (np.abs(w_o[:, 0]) > 0) & (k > 0)
    gpu_live = np.abs(f[:, 0]) > 0
    # same outcome (zero or not) for all but samples at a branch boundary
    # The sampler's own draws must agree sample for sample. A proposed normal reaches the branch
    # test ux <= F(wi.n) with ux computed from n in float32; in the grazing band both sides are
    # within a few float32 steps of 1 (1 - F ~ wi.n down to 1e-6), so rounding differences between
    # the GPU's ux and the oracle's flip a small share of outcomes and change the weight there.
    diff = (live != gpu_live) & same_n
    print(f"  outcome differs: {diff.mean():.2e} (by technique {[int(np.sum(diff & (tech == t))) for t in (0, 1, 2)]})")
    assert np.sum(diff & (tech == 0)) <= 2 and np.mean(diff[tech > 0]) < 5e-3
    both = live & gpu_live & same_n
    assert both.sum() > n_s // 2
    dwo = np.abs(wo - wo_o).max(1)
    rk = np.abs(k / np.where(factor > 0, factor, 1) - 1)
    est_o = w_o[:, 0].astype(np.float64) * factor
    rf = np.abs(f[:, 0] / np.where(est_o != 0, est_o, 1) - 1)
    for name, v, tol in (("wo", dwo, 2e-4), ("factor", rk, 2e-3), ("estimate", rf, 5e-3)):
        for label, sel, limit in (("sampler", both & (tech == 0), 2e-4), ("bands", both & (tech > 0), 5e-2)):
            frac = np.mean(v[sel] > tol)
            print(f"  {name} {label}: median {np.median(v[sel]):.2e} 99.9% {np.quantile(v[sel], 0.999):.2e} "
                  f"beyond {tol}: {frac:.2e}")
            assert frac < limit, (name, label, frac)
    # the reported pdf is the sampler's own
    rp = np.abs(pdf[both] / pdf_o[both] - 1)
    assert np.mean(rp[tech[both] == 0] > 1e-3) < 2e-4 and np.mean(rp > 1e-3) < 5e-2
    # every technique contributed comparable samples (none of the grazing band's for Beckmann
    # alpha 0.3: its grazing normals need a ux that rounds to 1 in float32, 1 - ux = exp(-tan^2/a^2))
    for t in ((0, 1, 2) if dist == rtg.DIST_GGX else (0, 1)):
        assert np.sum(both & (tech == t)) >= 5, t


def test_band_maps_have_the_densities_of_the_factor():
    """The proposals' maps from (ux, uy) to normals have the densities p1, p2 the MIS factor uses:
    in the parameterisation (c = wi.n, azimuth) the solid angle is dc dphi, so p(n) = p(c) / 2 pi
    when the azimuth is uniform and independent of c. Checked: the normals are unit vectors, the
    azimuth around wi is uniform (KS), and the counts of c in bins match the integrals of p1, p2
    over the bins (chi-square)."""
    from scipy import stats
    rng = np.random.default_rng(5)
    cos_c = np.sqrt(1 - (ETA_EXT / ETA_INT) ** 2)
    for cz in (-0.95, -0.8):
        wi = _wi(cz)
        m = 400_000
        ux, uy = rng.random(m), rng.random(m)
        for tech in (1, 2):
            n = band_normals(wi, np.full(m, tech), ux, uy, cos_c)
            assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-12)
            s, t = make_frame(wi)
            az = (np.arctan2(n @ t, n @ s) / (2 * np.pi)) % 1.0
            assert stats.kstest(az, "uniform").pvalue > 1e-3
            c = n @ wi
            if tech == 1:
                edges = np.concatenate([[0.0], np.geomspace(1e-6, B1, 24)])   # in x = -c - cos_c
                obs = np.histogram(-c - cos_c, edges)[0]
                expect = (np.sqrt(edges[1:]) - np.sqrt(edges[:-1])) / np.sqrt(B1)   # integral of 2 pi p1
            else:
                edges = np.concatenate([[0.0], np.geomspace(1e-8, G1, 24)])
                obs = np.histogram(c, edges)[0]
                expect = (np.sqrt(edges[1:]) - np.sqrt(edges[:-1])) / np.sqrt(G1)   # integral of 2 pi p2
            assert abs(expect.sum() - 1) < 1e-9 and obs.sum() == m
            chi2 = np.sum((obs - m * expect) ** 2 / (m * expect))
            assert chi2 < stats.chi2.ppf(0.999, len(obs) - 1), (tech, chi2)
            # the density formulas evaluated on the sampled normals agree with the bins' integrals
            p1, p2 = band_densities(wi, n, cos_c)
            dens = 2 * np.pi * (p1 if tech == 1 else p2)
            assert np.all(dens > 0)


@pytest.mark.parametrize("cz", [-0.95, -0.85, -0.78])
def test_exit_proposals_keep_the_mean_and_bound_the_weights(ctx, cz):
    mat = scenes.dielectric(0.1)
    m = 4_000_000
    rng = np.random.default_rng(int(-cz * 100))
    wi = np.tile(_wi(cz), (m, 1))
    L = lambda d: 1.0 + 0.8 * d[:, 0] + 0.3 * (d[:, 2] > 0)
    wo_p, f_p, _, _ = ctx.bsdf_sample(mat, wi, rng.random((m, 2)))
    plain = f_p[:, 0].astype(np.float64) * L(wo_p)
    wo_m, f_m, _, _ = ctx.exit_mis_sample(mat, wi, rng.random((m, 3)))
    mis = f_m[:, 0].astype(np.float64) * L(wo_m)
    se = np.sqrt(plain.var() / m + mis.var() / m)
    print(f"  cz {cz}: plain {plain.mean():.5f} (var {plain.var():.2f}, max {plain.max():.0f}) "
          f"exit MIS {mis.mean():.5f} (var {mis.var():.2f}, max {mis.max():.1f}) z {(mis.mean() - plain.mean()) / se:.2f}")
    assert abs(mis.mean() - plain.mean()) < 4.5 * se
    assert mis.max() < 100 and plain.max() > 300
    assert mis.var() < 0.5 * plain.var()


def test_exit_proposals_over_random_directions(ctx):
    """8M directions uniform over the denser side (random azimuth, including the rounding edge of
    the critical angle, where the sampler transmits with 1 - F of a few float32 steps): the same
    mean as the plain sampler, and no weight above 100 wherever the proposals apply (the plain
    sampler's reach 10^3-10^6 there)."""
    mat = scenes.dielectric(0.1)
    rng = np.random.default_rng(1)
    m = 8_000_000
    cz = -rng.random(m)
    ph = 2 * np.pi * rng.random(m)
    sz = np.sqrt(1 - cz * cz)
    wi = np.stack([sz * np.cos(ph), sz * np.sin(ph), cz], 1)
    L = lambda d: 1.0 + 0.8 * d[:, 0] - 0.5 * d[:, 1] + 0.3 * (d[:, 2] > 0)
    wo_p, f_p, _, _ = ctx.bsdf_sample(mat, wi, rng.random((m, 2)))
    plain = f_p[:, 0].astype(np.float64) * L(wo_p)
    wo_m, f_m, _, k = ctx.exit_mis_sample(mat, wi, rng.random((m, 3)))
    mis = f_m[:, 0].astype(np.float64) * L(wo_m)
    assert np.all(np.isfinite(mis))
    exit_ = cz < -np.sqrt(1 - (ETA_EXT / ETA_INT) ** 2)
    se = np.sqrt(plain.var() / m + mis.var() / m)
    print(f"  plain {plain.mean():.5f} (max {plain.max():.0f}, exit region {plain[exit_].max():.0f}) "
          f"exit MIS {mis.mean():.5f} (exit region max {mis[exit_].max():.1f}) z {(mis.mean() - plain.mean()) / se:.2f}")
    assert abs(mis.mean() - plain.mean()) < 4.5 * se
    assert mis[exit_].max() < 100
    assert np.all(k[~exit_ & (np.abs(f_m[:, 0]) > 0)] == 1.0)


@pytest.mark.parametrize("cz", [0.9, 0.3, -0.5])
def test_plain_sampler_runs_unchanged_elsewhere(ctx, cz):
    """Entering rays and rays beyond the macro-surface's critical angle: the plain sampler, bit for bit."""
    mat = scenes.dielectric(0.1)
    m = 100_000
    rng = np.random.default_rng(3)
    u = rng.random((m, 3)).astype(np.float32)
    wi = np.tile(_wi(cz), (m, 1))
    wo_m, f_m, pdf_m, k = ctx.exit_mis_sample(mat, wi, u)
    wo_p, f_p, pdf_p, _ = ctx.bsdf_sample(mat, wi, u[:, 1:])
    live = np.abs(f_p[:, 0]) > 0
    assert np.array_equal(wo_m[live], wo_p[live]) and np.array_equal(f_m, f_p) and np.array_equal(pdf_m[live], pdf_p[live])
    assert np.all(k[live] == 1.0)


def _block(img, b=32):
    h, w, _ = img.shape
    return img.reshape(h // b, b, w // b, b, 3).mean((1, 3))


@pytest.mark.parametrize("mode", [rtg.GUIDING_EXIT, rtg.GUIDING_LEARNED])
def test_guided_modes_match_the_converged_reference(ctx, mode):
    """512x512 glossy-indirect renders (8 x 512 spp) against the committed 2^20-spp plain reference:
    the image mean per channel agrees within the renders' standard error, and per 32x32 block the
    mean over the renders agrees within the renders' spread and the reference's own noise (a t
    statistic with 7 degrees of freedom: mean square about 7/5, few beyond 4)."""
    z = np.load(os.path.join(HERE, "golden", "glossy_cornell_ref.npz"))
    ref = z["mean"].astype(np.float64)
    ref_var = z["var_mean"].astype(np.float64)
    sd = scenes.glossy_cornell()
    g = ctx.upload(sd)
    blocks, means = [], []
    for s in range(8):
        out, st = g.render(sd.camera, dict(sd.render, spp=512, seed=300 + s, guiding=mode))
        assert st.nan_samples == 0
        img = rtg.resolve(out).astype(np.float64)
        blocks.append(_block(img))
        means.append(img.reshape(-1, 3).mean(0))
    M = np.stack(means)
    sem = M.std(0, ddof=1) / np.sqrt(len(M))
    rm = ref.reshape(-1, 3).mean(0)
    print(f"  mode {mode}: image mean {M.mean(0).round(6)} (sem {sem.round(6)}) reference {rm.round(6)}")
    assert np.all(np.abs(M.mean(0) - rm) < 5 * sem + 1e-5 * rm)
    B = np.stack(blocks)
    rb = _block(ref)
    rvar = _block(ref_var) / (32 * 32)  # block mean of independent pixel noise
    se = np.sqrt(B.var(0, ddof=1) / len(B) + rvar) + 1e-5 * np.abs(rb)
    zz = (B.mean(0) - rb) / se
    print(f"  blocks: max|z| {np.abs(zz).max():.2f} mean z^2 {np.mean(zz * zz):.2f} beyond 4: {np.mean(np.abs(zz) > 4):.4f}")
    assert np.mean(zz * zz) < 2.5 and np.mean(np.abs(zz) > 4) < 0.03 and np.abs(zz).max() < 12
