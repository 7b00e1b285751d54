"""2-D transforms, linear steppers and polynomial reactions on the GPU.

- ``fft`` / ``ifft`` in 2-D (rfftn / irfftn over the last two axes,
  _spectral.py:614-721) against numpy, including spectra whose DC and Nyquist
  columns carry imaginary parts that irfftn ignores;
- the linear steppers in 2-D (ETDRK0; complex exp(L dt) for advection and
  dispersion) against the FP32 oracle at steps 1 and 10;
- the polynomial reaction steppers in 2-D (AllenCahn, FisherKPP,
  SwiftHohenberg; ETDRK orders 1-4) and a polynomial stepper with a complex
  linear operator (odd derivative terms) against the FP32 oracle.
"""
import numpy as np
import pytest

import exponax_amd as ex
from oracle import stepper as ost
from tests.helpers import RTOL_STEP, ics, max_rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [32, 64, 128, 256, 512])
def test_fft2d_against_numpy(device, N):
    import torch

    print(f"[case] fft2d N={N}", flush=True)
    rng = np.random.default_rng(N)
    u = rng.standard_normal((2, 3, N, N)).astype(np.float32)
    uh = ex.spectral.fft(torch.tensor(u, device=device), num_spatial_dims=2)
    want = np.fft.rfftn(u.astype(np.float64), axes=(-2, -1))
    err = np.abs(uh.cpu().numpy() - want).max() / np.abs(want).max()
    back = ex.spectral.ifft(uh, num_spatial_dims=2).cpu().numpy()
    # a general spectrum: irfftn drops Im of the last axis' DC and Nyquist columns
    z = (rng.standard_normal((2, N, N // 2 + 1)) + 1j * rng.standard_normal((2, N, N // 2 + 1))).astype(np.complex64)
    got_i = ex.spectral.ifft(torch.tensor(z, device=device), num_spatial_dims=2).cpu().numpy()
    want_i = np.fft.irfftn(z.astype(np.complex128), s=(N, N), axes=(-2, -1))
    err_i = np.abs(got_i - want_i).max() / np.abs(want_i).max()
    print(f"  rfftn err {err:.2e}, roundtrip {np.abs(back - u).max():.2e}, irfftn err {err_i:.2e}", flush=True)
    assert err <= 2e-6 and err_i <= 2e-6
    np.testing.assert_allclose(back, u, atol=2e-5)


def _linear_pair(name, N):
    L = 2.0
    if name == "advection":
        kw = dict(velocity=np.array([0.8, -0.3]))
        return ex.stepper.Advection(2, L, N, 0.01, **kw), ost.Advection(2, L, N, 0.01, dtype=np.float32, **kw)
    if name == "advection_diffusion":
        kw = dict(velocity=-0.5, diffusivity=0.01)
        return (ex.stepper.AdvectionDiffusion(2, L, N, 0.01, **kw),
                ost.AdvectionDiffusion(2, L, N, 0.01, dtype=np.float32, **kw))
    if name == "dispersion":
        return (ex.stepper.Dispersion(2, L, N, 1e-4, dispersivity=0.2),
                ost.Dispersion(2, L, N, 1e-4, dispersivity=0.2, dtype=np.float32))
    if name == "diffusion":
        return (ex.stepper.Diffusion(2, L, N, 0.01, diffusivity=0.02),
                ost.Diffusion(2, L, N, 0.01, diffusivity=0.02, dtype=np.float32))
    return (ex.stepper.HyperDiffusion(2, L, N, 0.01, hyper_diffusivity=1e-4),
            ost.HyperDiffusion(2, L, N, 0.01, hyper_diffusivity=1e-4, dtype=np.float32))


def _check(gpu, cpu, u0, device, label):
    import torch

    x = torch.tensor(u0, device=device)
    traj = gpu.rollout_batch(x, 10).cpu().numpy()
    ug, uc = x, u0
    for s in range(1, 11):
        ug = gpu.step(ug)
        uc = cpu.step_batch(uc)
        if s in (1, 10):
            e_step = max_rel_l2(ug.cpu().numpy(), uc)
            e_roll = max_rel_l2(traj[:, s - 1], uc)
            print(f"  {label} step {s}: steps {e_step:.2e}, rollout {e_roll:.2e}", flush=True)
            assert e_step <= RTOL_STEP and e_roll <= RTOL_STEP


@pytest.mark.parametrize("N", [64, 128])
@pytest.mark.parametrize("name", ["advection", "advection_diffusion", "dispersion", "diffusion", "hyper_diffusion"])
def test_linear_2d_parity(device, name, N):
    print(f"[case] 2-D {name} N={N}", flush=True)
    gpu, cpu = _linear_pair(name, N)
    u0, _ = ics(3, 2, 2.0, N, 5, seed=61)
    _check(gpu, cpu, u0, device, name)


def _reaction_pair(name, N, order):
    if name == "allen_cahn":
        kw = dict(diffusivity=5e-3, order=order)
        return (ex.stepper.reaction.AllenCahn(2, 1.0, N, 1e-3, **kw),
                ost.AllenCahn(2, 1.0, N, 1e-3, dtype=np.float32, **kw), 1.0)
    if name == "fisher_kpp":
        kw = dict(diffusivity=0.01, reactivity=1.0, order=order)
        return (ex.stepper.reaction.FisherKPP(2, 1.0, N, 1e-3, **kw),
                ost.FisherKPP(2, 1.0, N, 1e-3, dtype=np.float32, **kw), 1.0)
    kw = dict(order=order)
    return (ex.stepper.reaction.SwiftHohenberg(2, 20.0, N, 0.05, **kw),
            ost.SwiftHohenberg(2, 20.0, N, 0.05, dtype=np.float32, **kw), 20.0)


@pytest.mark.parametrize("N", [64, 256])
@pytest.mark.parametrize("name,order", [("allen_cahn", 2), ("fisher_kpp", 2), ("swift_hohenberg", 2),
                                        ("allen_cahn", 1), ("allen_cahn", 3), ("swift_hohenberg", 4)])
def test_reaction_2d_parity(device, name, order, N):
    print(f"[case] 2-D {name} order {order} N={N}", flush=True)
    gpu, cpu, L = _reaction_pair(name, N, order)
    u0, _ = ics(3, 2, L, N, 4, seed=62)
    if name == "fisher_kpp":
        u0 = (0.5 + 0.4 * u0 / np.abs(u0).max()).astype(np.float32)  # a population in (0, 1)
    _check(gpu, cpu, u0, device, name)


class _OraclePolynomial(ost._PolynomialReaction):
    """Oracle counterpart of GeneralPolynomialStepper: L = sum_j a_j sum_d
    (i kappa_d)^j (stepper/generic/_polynomial.py:11-110), N = F[sum c_i u^i]."""

    def __init__(self, D, L, N, dt, lin, poly, order):
        self.lin, self.pc, self.dealiasing_fraction = tuple(lin), list(poly), 2 / 3
        self._kw_dtype = np.float32
        super().__init__(D, L, N, dt, num_channels=1, order=order, num_circle_points=16, circle_radius=1.0,
                         dtype=np.float32)

    def _build_linear_operator(self, d):
        out = np.zeros((1,) + d.shape[1:], dtype=np.complex128)
        for j, a in enumerate(self.lin):
            out = out + a * np.sum(d**j, axis=0, keepdims=True)
        return out

    def _poly(self):
        return self.pc


@pytest.mark.parametrize("order", [1, 2])
def test_polynomial_2d_complex_linear(device, order):
    print(f"[case] 2-D polynomial, complex L, order {order}", flush=True)
    lin, poly, N, L, dt = (0.5, -0.4, 0.01), (0.0, 0.0, -0.5), 64, 2.0, 1e-3
    gpu = ex.stepper.generic.GeneralPolynomialStepper(2, L, N, dt, linear_coefficients=lin,
                                                      polynomial_coefficients=poly, order=order)
    cpu = _OraclePolynomial(2, L, N, dt, lin, poly, order)
    u0, _ = ics(3, 2, L, N, 4, seed=63)
    _check(gpu, cpu, u0, device, "polynomial")


@pytest.mark.parametrize("D,N,order", [(1, 64, 2), (1, 256, 2), (1, 128, 4), (2, 64, 2), (2, 128, 1)])
def test_cahn_hilliard_parity(device, D, N, order):
    """stepper/reaction/_cahn_hilliard.py: the Laplacian of the cubic (a
    polynomial whose transform the kernels multiply by scale * Laplace)."""
    print(f"[case] Cahn-Hilliard D={D} N={N} order {order}", flush=True)
    kw = dict(diffusivity=1e-2, gamma=1e-3, order=order)
    gpu = ex.stepper.reaction.CahnHilliard(D, 2.0, N, 1e-3, **kw)
    cpu = ost.CahnHilliard(D, 2.0, N, 1e-3, dtype=np.float32, **kw)
    u0, _ = ics(3, D, 2.0, N, 4, seed=64)
    _check(gpu, cpu, u0, device, "cahn_hilliard")
