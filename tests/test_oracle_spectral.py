"""Pins of the oracle's spectral layer (oracle/spectral.py) on the CPU.

Each check comes from the reference's own tests (exponax tests/...) or from
what the mathematics fixes (brute-force DFT, closed-form derivatives,
Parseval), chosen so that a dropped term, a wrong sign or a transposed axis
fails.
"""
import json
import os

import numpy as np
import pytest

from oracle import spectral as sp

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_filter_masks_match_reference_literals():
    """The literal masks of exponax tests/test_filter_masks.py:6-451."""
    with open(os.path.join(GOLDEN, "filter_masks.json")) as f:
        cases = json.load(f)["cases"]
    assert len(cases) == 14
    for c in cases:
        fn = getattr(sp, c["function"])
        got = fn(*c["args"], **c["kwargs"])
        np.testing.assert_array_equal(got, np.array(c["expected"], dtype=bool), err_msg=str(c["args"]))


def test_shape_utilities():
    """exponax tests/test_shape_utilties.py:4-23."""
    assert sp.spatial_shape(2, 64) == (64, 64)
    assert sp.wavenumber_shape(1, 64) == (33,)
    assert sp.wavenumber_shape(2, 64) == (64, 33)
    assert sp.wavenumber_shape(3, 65) == (65, 65, 33)


def test_modes_slices():
    """exponax tests/test_mode_slices_generation.py:4-40."""
    assert sp.get_modes_slices(1, 10) == ((slice(None), slice(None, 6)),)
    assert sp.get_modes_slices(2, 10) == (
        (slice(None), slice(None, 5), slice(None, 6)),
        (slice(None), slice(-5, None), slice(None, 6)),
    )
    assert sp.get_modes_slices(2, 11) == (
        (slice(None), slice(None, 6), slice(None, 6)),
        (slice(None), slice(-5, None), slice(None, 6)),
    )


def test_wavenumbers_closed_form():
    wn = sp.build_wavenumbers(2, 8)
    assert wn.shape == (2, 8, 5)
    np.testing.assert_array_equal(wn[0][:, 0], [0, 1, 2, 3, -4, -3, -2, -1])
    np.testing.assert_array_equal(wn[1][0], [0, 1, 2, 3, 4])
    d = sp.build_derivative_operator(1, 3.0, 16)
    np.testing.assert_allclose(d[0], 1j * 2 * np.pi / 3.0 * np.arange(9))


def _dft_r2c(x):
    """Brute-force O(N^2) real DFT along the last axis (norm="backward")."""
    n = x.shape[-1]
    k = np.arange(n // 2 + 1)[:, None]
    m = np.arange(n)[None, :]
    return x @ np.exp(-2j * np.pi * k * m / n).T


@pytest.mark.parametrize("n", [8, 10, 16, 33])
def test_fft_matches_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n))
    np.testing.assert_allclose(sp.fft(x, num_spatial_dims=1), _dft_r2c(x), atol=1e-12)
    # 2-D: c2c on the leading axis, r2c on the last
    x2 = rng.standard_normal((1, n, n))
    ref = np.fft.fft(_dft_r2c(x2), axis=-2)
    np.testing.assert_allclose(sp.fft(x2, num_spatial_dims=2), ref, atol=1e-10)


def test_ifft_conventions():
    """irfft ignores Im of the DC and Nyquist bins, scales by 1/N."""
    rng = np.random.default_rng(1)
    n = 16
    x = rng.standard_normal((1, n))
    X = sp.fft(x, num_spatial_dims=1)
    np.testing.assert_allclose(sp.ifft(X, num_spatial_dims=1, num_points=n), x, atol=1e-13)
    X2 = X.copy()
    X2[0, 0] += 5j
    X2[0, -1] -= 3j
    np.testing.assert_allclose(sp.ifft(X2, num_spatial_dims=1, num_points=n), x, atol=1e-13)
    with pytest.raises(ValueError):
        sp.ifft(X)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_roundtrip(d):
    """exponax tests/test_spectral.py:18-31."""
    g = sp.make_grid(d, 5.0, 32)
    u = np.sin(2 * np.pi * g[0:1] / 5.0)
    np.testing.assert_allclose(sp.ifft(sp.fft(u, num_spatial_dims=d), num_spatial_dims=d, num_points=32), u,
                               atol=1e-5)


@pytest.mark.parametrize("d,axis", [(1, 0), (2, 0), (2, 1)])
def test_derivative_of_sine(d, axis):
    """exponax tests/test_spectral.py:42-69."""
    L, k = 3.0, 3
    g = sp.make_grid(d, L, 64)
    u = np.sin(k * 2 * np.pi * g[axis:axis + 1] / L)
    du = sp.derivative(u, L, order=1)
    expected = k * 2 * np.pi / L * np.cos(k * 2 * np.pi * g[axis:axis + 1] / L)
    np.testing.assert_allclose(du[axis], expected[0], atol=1e-4)


def test_laplace_and_biharmonic_eigenvalues():
    """exponax tests/test_spectral.py:87-99, 148-157: sum_d (i kappa_d)^p."""
    d = sp.build_derivative_operator(2, 4.0, 16)
    kap = sp.build_scaled_wavenumbers(2, 4.0, 16)
    np.testing.assert_allclose(sp.build_laplace_operator(d), -(kap[0:1] ** 2 + kap[1:2] ** 2), atol=1e-10)
    np.testing.assert_allclose(sp.build_laplace_operator(d, order=4), kap[0:1] ** 4 + kap[1:2] ** 4, atol=1e-8)
    np.testing.assert_allclose(sp.build_laplace_operator(d, order=0), np.ones((1, 16, 9)))
    with pytest.raises(ValueError):
        sp.build_laplace_operator(d, order=3)


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("n", [10, 11])
def test_scaling_array_norm_compensation(d, n):
    """exponax tests/test_spectral_scaling_arrays.py:8-36."""
    x = np.random.default_rng(0).standard_normal((1,) + (n,) * d)
    back = np.fft.rfftn(x, axes=tuple(range(-d, 0)))
    fwd = np.fft.rfftn(x, axes=tuple(range(-d, 0)), norm="forward")
    np.testing.assert_allclose(back / sp.build_scaling_array(d, n, mode="norm_compensation"), fwd, atol=1e-12)


def test_dealias_cutoffs():
    """nonlin_fun/_base.py:59-71: keep |k| <= 2/3 (N//2) - 1."""
    for n, kmax in [(64, 20), (128, 41), (256, 84), (512, 169), (1024, 340)]:
        m = sp.dealiasing_mask(1, n, 2 / 3)[0]
        assert m.sum() == kmax + 1 and m[kmax] and not m[kmax + 1]
    m2 = sp.dealiasing_mask(2, 128, 2 / 3)[0]
    assert m2.shape == (128, 65) and m2.sum() == (2 * 41 + 1) * 42  # hypercube, not a disk
    assert sp.dealiasing_mask(1, 64, None) is None


def test_spectrum_amplitude_and_parseval():
    """exponax tests/test_spectrum.py:12-30 and the Parseval identity of
    :163-216 (sum of the power spectrum = mean(u^2) / 2)."""
    g = sp.make_grid(1, 2 * np.pi, 128)
    assert sp.get_spectrum(3.0 * np.sin(g), power=False)[0, 1] == pytest.approx(3.0)
    assert sp.get_spectrum(3.0 * np.sin(3 * g) + 4.0 * np.cos(3 * g), power=False)[0, 3] == pytest.approx(5.0)
    assert sp.get_spectrum(3.0 * np.ones_like(g), power=False)[0, 0] == pytest.approx(3.0)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((1, 64))
    assert sp.get_spectrum(u).sum() == pytest.approx(0.5 * np.mean(u**2))
    u2 = rng.standard_normal((1, 32, 32))
    g2 = sp.make_grid(2, 2 * np.pi, 32)
    assert sp.get_spectrum(3.0 * np.sin(g2[0:1]), power=False)[0, 1] == pytest.approx(3.0)
    assert sp.get_spectrum(u2).sum() == pytest.approx(0.5 * np.mean(u2**2), rel=0.2)


@pytest.mark.parametrize("d", [1, 2])
def test_map_between_resolutions_roundtrip(d):
    """exponax tests/test_interpolation.py:326-353 on a band-limited field."""
    from oracle import ic

    u = ic.tfs(d, 1.0, 32, 4, np.random.default_rng(0))
    down = sp.map_between_resolutions(u, 24)
    np.testing.assert_allclose(sp.map_between_resolutions(down, 32), u, atol=3e-5)
    up = sp.map_between_resolutions(u, 48)
    np.testing.assert_allclose(sp.map_between_resolutions(up, 32), u, atol=3e-5)
