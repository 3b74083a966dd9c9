"""Spectral IOR data to RGB (rtg_spectrum_to_rgb; core/Color.cpp:423-443, Color.h:196-248), the
conversion the scene loader applies to a named metal's data/IOR spectra. Host code only: pinned by
the CIE 1931 2-degree observer's published values and integrals and by the standard sRGB primaries
matrix (IEC 61966-2-1), which inverts the XYZ -> RGB step."""
import numpy as np
import pytest

from raytracer_amd import rtg

# CIE 1931 2-degree colour-matching functions (x, y, z) at a few wavelengths, and their sums over
# the 1 nm table from 360 to 830 nm
CMF = {450: (0.3362, 0.038, 1.77211), 555: (0.5120501, 1.0, 0.005749999), 600: (1.0622, 0.631, 0.0008)}
SUMS = (106.8655, 106.8570, 106.8923)
# linear sRGB -> XYZ (D65)
RGB2XYZ = np.array([[0.412453, 0.357580, 0.180423],
                    [0.212671, 0.715160, 0.072169],
                    [0.019334, 0.119193, 0.950227]])
SCALE = 470.0 / (106.856895 * 471)


def _rgb(lam, val):
    return rtg.spectrum_to_rgb(np.asarray(lam, np.float32), np.asarray(val, np.float32)).astype(np.float64)


def _peaks(weights):
    """Narrow triangles of height w at integer wavelengths: the table is read at w there and at 0
    at every other nanometre."""
    lam, val = [355.0], [0.0]
    for l, w in sorted(weights.items()):
        lam += [l - 1, l, l + 1]
        val += [0.0, w, 0.0]
    return lam + [835.0], val + [0.0]


def test_constant_spectrum():
    """A constant c integrates to c times the sums of the matching functions: Y = c (470/471)
    (106.8570/106.856895), X/Y and Z/Y the ratios of the sums (within the 4-digit sums)."""
    for c in (1.0, 0.25, 3.5):
        rgb = _rgb([400.0, 700.0], [c, c])
        assert np.all(rgb > 0)
        xyz = RGB2XYZ @ rgb
        want = c * SCALE * np.array(SUMS)
        assert np.allclose(xyz, want, rtol=2e-5), (xyz, want)


def test_cie_table_alignment_and_matrix():
    """Three peaks (450, 555, 600 nm) whose mixture lies inside the sRGB gamut: the RGB maps back
    to the weighted sum of the published values at exactly those wavelengths (a table shifted by one
    nanometre, a transposed matrix or a wrong scale fails)."""
    w = {450: 0.5, 555: 1.0, 600: 1.0}
    rgb = _rgb(*_peaks(w))
    assert np.all(rgb > 0)
    want = SCALE * sum(wt * np.array(CMF[l]) for l, wt in w.items())
    assert np.allclose(RGB2XYZ @ rgb, want, rtol=1e-4), (RGB2XYZ @ rgb, want)
    # a shifted table would read the neighbours
    assert not np.allclose(RGB2XYZ @ _rgb(*_peaks({451: 0.5, 556: 1.0, 601: 1.0})), want, rtol=1e-3)


def test_negative_channels_clamp():
    """Monochromatic 555 nm lies outside the sRGB gamut: its blue channel is negative and clamps to
    zero; red and green follow from the XYZ -> RGB matrix (the inverse of the primaries matrix)."""
    rgb = _rgb(*_peaks({555: 1.0}))
    xyz = SCALE * np.array(CMF[555])
    m = np.linalg.inv(RGB2XYZ) @ xyz
    assert m[2] < 0 and rgb[2] == 0.0
    assert np.allclose(rgb[:2], m[:2], rtol=1e-4)


def test_interpolation_and_extension():
    """Linear between samples and constant beyond the ends (Color.h:239-248): a ramp given on
    [500, 600] by 5 samples equals the same function tabulated at every nanometre of 360-830; a
    pair (0, 0) ahead of data starting at 400 nm (what a text header line becomes) makes the
    spectrum below 400 nm the line from 0 to the first value."""
    f = lambda l: 0.3 + 0.01 * (l - 500.0)
    coarse = np.linspace(500, 600, 5)
    fine = np.arange(360, 831, dtype=np.float64)
    full = f(np.clip(fine, 500, 600))
    assert np.allclose(_rgb(coarse, f(coarse)), _rgb(fine, full), rtol=1e-5)
    lam = np.array([0.0, 400.0, 700.0])
    val = np.array([0.0, 2.0, 1.0])
    g = np.interp(fine, lam, val)
    assert np.allclose(_rgb(lam, val), _rgb(fine, g), rtol=1e-5)


def test_rejects_short_spectra():
    with pytest.raises(rtg.RtgError):
        rtg.spectrum_to_rgb(np.array([500.0], np.float32), np.array([1.0], np.float32))
