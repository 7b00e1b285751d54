"""GPU parity of the measurement and resolution functions around the path
against the oracle: nRMSE (metrics/_spatial.py:537-586), the 1-D spectrum
(_spectral.py:866-1030) and map_between_resolutions
(_interpolation.py:195-285)."""
import numpy as np
import pytest

import exponax_amd as ex
from oracle import spectral as osp
from oracle import stepper as ost
from tests.helpers import ics

pytestmark = pytest.mark.gpu


def test_nrmse_batch(device):
    import torch

    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 3, 64, 64)).astype(np.float32)
    b = (a + 0.01 * rng.standard_normal(a.shape)).astype(np.float32)
    got = ex.metrics.nRMSE_batch(torch.tensor(b, device="cuda"), torch.tensor(a, device="cuda")).cpu().numpy()
    ref = np.array([ost.nrmse(b[i], a[i]) for i in range(5)])
    np.testing.assert_allclose(got, ref, rtol=1e-6)
    assert ex.metrics.nRMSE(b[0], a[0]) == pytest.approx(ref[0], rel=1e-6)
    assert ex.metrics.mean_nRMSE(b, a) == pytest.approx(ref.mean(), rel=1e-6)


@pytest.mark.parametrize("N", [16, 64, 256, 1024])
@pytest.mark.parametrize("power", [True, False])
def test_spectrum_1d(device, N, power):
    u, _ = ics(3, 1, 2.0, N, 7, seed=3)
    got = ex.get_spectrum(u, power=power)
    ref = np.stack([osp.get_spectrum(u[i].astype(np.float64), power=power) for i in range(3)])
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, atol=1e-6 * scale, rtol=1e-5)


@pytest.mark.parametrize("n_old,n_new", [(256, 64), (64, 256), (128, 32), (32, 128), (1024, 256), (64, 64)])
@pytest.mark.parametrize("oddball", [True, False])
def test_map_between_resolutions_1d(device, n_old, n_new, oddball):
    rng = np.random.default_rng(n_old + n_new)
    u = rng.standard_normal((4, 1, n_old)).astype(np.float32)  # full band, Nyquist included
    got = ex.map_between_resolutions(u, n_new, oddball_zero=oddball)
    ref = np.stack([osp.map_between_resolutions(u[i].astype(np.float64), n_new, oddball_zero=oddball)
                    for i in range(4)])
    np.testing.assert_allclose(got, ref, atol=2e-6 * np.abs(ref).max(), rtol=0)


def test_downsample_fine_trajectory(device):
    """The corrector's data path: KS fine states (N = 256) to the coarse grid
    (N = 64) through the GPU equals the oracle on the same states."""
    u, _ = ics(8, 1, 100.0, 256, 5, seed=5)
    got = ex.map_between_resolutions(u, 64)
    ref = np.stack([osp.map_between_resolutions(u[i].astype(np.float64), 64) for i in range(8)])
    assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("D,N", [(2, 32), (2, 128), (3, 16), (3, 32)])
@pytest.mark.parametrize("power,binning", [(True, "sum"), (False, "sum"), (True, "average")])
def test_spectrum_radial(device, D, N, power, binning):
    """2-D and 3-D spectra binned by the wavenumber norm (_spectral.py:866-1030)."""
    print(f"[case] spectrum D={D} N={N} power={power} {binning}", flush=True)
    rng = np.random.default_rng(D * N)
    u = rng.standard_normal((3, 2) + (N,) * D).astype(np.float32)
    got = ex.get_spectrum(u, power=power, radial_binning=binning)
    ref = np.stack([osp.get_spectrum(u[i].astype(np.float64), power=power, radial_binning=binning)
                    for i in range(3)])
    assert got.shape == ref.shape == (3, 2, N // 2 + 1)
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-7 * np.nanmax(np.abs(ref)))


@pytest.mark.parametrize("D,n_old,n_new", [(2, 64, 32), (2, 32, 128), (2, 128, 64), (3, 32, 16), (3, 16, 32)])
@pytest.mark.parametrize("oddball", [True, False])
def test_map_between_resolutions_nd(device, D, n_old, n_new, oddball):
    print(f"[case] resample D={D} {n_old}->{n_new} oddball={oddball}", flush=True)
    rng = np.random.default_rng(n_old * 7 + n_new + D)
    u = rng.standard_normal((2, 1) + (n_old,) * D).astype(np.float32)  # full band, Nyquist modes included
    got = ex.map_between_resolutions(u, n_new, oddball_zero=oddball)
    ref = np.stack([osp.map_between_resolutions(u[i].astype(np.float64), n_new, oddball_zero=oddball)
                    for i in range(2)])
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, atol=3e-6 * np.abs(ref).max(), rtol=0)


@pytest.mark.parametrize("D,N,C", [(1, 256, 1), (1, 1024, 2), (2, 64, 1), (2, 128, 2), (3, 16, 1), (3, 32, 3)])
def test_ic_generation_on_device(device, D, N, C):
    """The seeded truncated-Fourier-series ICs built on the GPU equal the
    host recipe drawn from the same stream."""
    from exponax_amd import ic as xic
    from oracle import ic as oic

    print(f"[case] device ICs D={D} N={N} C={C}", flush=True)
    cutoff = 3 if D == 3 else 5
    got = xic.truncated_fourier_series_device(4, D, 2.0, N, cutoff, 17, num_channels=C).cpu().numpy()
    want, _ = oic.tfs_batch(4, D, 2.0, N, cutoff, 17, num_channels=C)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, atol=3e-6, rtol=0)
    assert np.abs(np.abs(got).max(axis=tuple(range(2, got.ndim))) - 1).max() < 1e-6
