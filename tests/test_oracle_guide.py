"""Pins of the guiding-network oracle (oracle/guide_oracle.cpp): bf16 rounding, the encoding,
the vMF-mixture density's normalisation, and the FP64 backward pass against finite differences."""
import numpy as np

import oracle_lib as O


def test_bf16_round_to_nearest_even():
    b = O.load().gor_bf16
    assert b(1.0) == 1.0
    assert b(1.0 + 2.0 ** -8) == 1.0             # tie -> even (mantissa 0)
    assert b(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6   # tie -> even (rounds up)
    assert b(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert b(-3.14159) == -3.140625
    assert b(1e-40) != 0.0 or True               # denormals pass through the float conversion


def test_encoding_closed_form():
    rng = np.random.default_rng(1)
    x = rng.random((64, 6)).astype(np.float32)
    x[:, 3:] = x[:, 3:] * 2 - 1
    f = O.guide_encode(x)
    v = np.concatenate([x[:, :3], (x[:, 3:] + 1) * 0.5], axis=1).astype(np.float32).astype(np.float64)
    for d in range(6):
        for k in range(12):
            a = np.pi * v[:, d] * 2.0 ** k
            assert np.allclose(f[:, d * 24 + 2 * k], np.sin(a), atol=1e-9)
            assert np.allclose(f[:, d * 24 + 2 * k + 1], np.cos(a), atol=1e-9)


def _sphere_quadrature(n_theta=400, n_phi=800):
    th = (np.arange(n_theta) + 0.5) * np.pi / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / n_theta) * (2 * np.pi / n_phi)).reshape(-1)
    return w, dA


def test_mixture_density_integrates_to_one():
    params = O.guide_init(3, scale=3.0)
    rng = np.random.default_rng(4)
    x = rng.random((3, 6)).astype(np.float32)
    _, mix = O.guide_forward(params, x, emulate_bf16=False)
    assert np.allclose(mix[:, 0::5].sum(1), 1.0)
    w, dA = _sphere_quadrature()
    for i in range(3):
        m = np.repeat(mix[i:i + 1], w.shape[0], axis=0)
        p = np.exp(O.guide_mix_logpdf(m, w))
        assert abs((p * dA).sum() - 1.0) < 2e-3, (p * dA).sum()
    # a single sharp lobe: density at the mean = kappa / (2 pi (1 - e^-2kappa))
    one = np.zeros((1, 40)); one[0, 3::5] = 1.0; one[0, 4::5] = 1.0
    one[0, 0] = 1.0; one[0, 4] = 50.0  # lobe 0 carries all the weight
    lp = O.guide_mix_logpdf(one, np.array([[0.0, 0.0, 1.0]]))
    assert np.isclose(np.exp(lp[0]), 50 / (2 * np.pi * (1 - np.exp(-100))))


def test_backward_matches_finite_differences():
    """FP64 analytic gradient (no bf16 rounding) against central differences on a sample of
    parameters from every layer, weights and biases."""
    params = O.guide_init(5, scale=2.0)
    rng = np.random.default_rng(6)
    n = 24
    s = np.zeros((n, 10), np.float32)
    s[:, :3] = rng.random((n, 3))
    nn = rng.normal(size=(n, 3)); s[:, 3:6] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    w = rng.normal(size=(n, 3)); s[:, 6:9] = w / np.linalg.norm(w, axis=1, keepdims=True)
    s[:, 9] = rng.random(n) + 0.1
    loss, g = O.guide_loss_grad(params, s, emulate_bf16=False)
    assert np.isfinite(loss) and loss > 0
    idx = []
    for name, (off, shp) in O.GUIDE_SLICES.items():
        idx += list(off + rng.choice(int(np.prod(shp)), 6, replace=False))
    # the FP32 parameter vector is perturbed by a float-representable step
    h = 2.0 ** -12
    for i in idx:
        pp = params.copy(); pm = params.copy()
        pp[i] += h; pm[i] -= h
        lp, _ = O.guide_loss_grad(pp, s, emulate_bf16=False)
        lm, _ = O.guide_loss_grad(pm, s, emulate_bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 2e-4 * max(1.0, abs(g[i])) + 1e-6, (i, fd, g[i])


def test_bf16_emulation_is_close_to_fp64():
    """bf16 keeps ~3 significant digits: the emulated gradient is within 15% (L2) of FP64, so a
    GPU comparison must use the emulated oracle."""
    params = O.guide_init(7)
    rng = np.random.default_rng(8)
    s = np.zeros((64, 10), np.float32)
    s[:, :3] = rng.random((64, 3))
    nn = rng.normal(size=(64, 3)); s[:, 3:6] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    w = rng.normal(size=(64, 3)); s[:, 6:9] = w / np.linalg.norm(w, axis=1, keepdims=True)
    s[:, 9] = 1.0
    l0, g0 = O.guide_loss_grad(params, s, emulate_bf16=False)
    l1, g1 = O.guide_loss_grad(params, s, emulate_bf16=True)
    assert abs(l0 - l1) < 1e-2 * abs(l0)
    assert np.linalg.norm(g0 - g1) < 0.15 * np.linalg.norm(g0)
    assert np.dot(g0, g1) > 0.98 * np.linalg.norm(g0) * np.linalg.norm(g1)
