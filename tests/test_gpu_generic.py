"""Generic stepper families on the GPU, by the equivalences the reference's
own tests state (tests/test_builtin_solvers.py): a generic stepper with the
matching coefficients reproduces each specific stepper, and a normalized
stepper reproduces the regular one (rel 1e-4 there; 1e-5 here)."""
import numpy as np
import pytest

import exponax_amd as ex
from exponax_amd.stepper import generic as g
from tests.helpers import ics, max_rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _one(stepper, u0):
    import torch

    return stepper.step(torch.tensor(u0, device="cuda")).cpu().numpy()


LINEAR = [
    (lambda: ex.stepper.Advection(1, 3.0, 64, 0.1, velocity=1.0), [0.0, -1.0]),
    (lambda: ex.stepper.Diffusion(1, 3.0, 64, 0.1, diffusivity=0.01), [0.0, 0.0, 0.01]),
    (lambda: ex.stepper.AdvectionDiffusion(1, 3.0, 64, 0.1, velocity=1.0, diffusivity=0.01), [0.0, -1.0, 0.01]),
    (lambda: ex.stepper.Dispersion(1, 3.0, 64, 0.1, dispersivity=0.01), [0.0, 0.0, 0.0, 0.01]),
    (lambda: ex.stepper.HyperDiffusion(1, 3.0, 64, 0.1, hyper_diffusivity=0.0001), [0.0, 0.0, 0.0, 0.0, -0.0001]),
]


@pytest.mark.parametrize("idx", range(len(LINEAR)))
def test_general_linear_matches_specific(device, idx):
    make, coefs = LINEAR[idx]
    spec = make()
    gen = g.GeneralLinearStepper(1, spec.domain_extent, spec.num_points, spec.dt, linear_coefficients=coefs)
    u0, _ = ics(3, 1, spec.domain_extent, spec.num_points, 5, seed=61)
    assert max_rel_l2(_one(gen, u0), _one(spec, u0)) <= TOL


def test_general_convection_and_gradient_norm_match_specific(device):
    u0, _ = ics(3, 1, 3.0, 64, 5, seed=62)
    burgers = ex.stepper.Burgers(1, 3.0, 64, 0.1, diffusivity=0.01)
    gen = g.GeneralConvectionStepper(1, 3.0, 64, 0.1, linear_coefficients=[0.0, 0.0, 0.01], convection_scale=1.0)
    assert max_rel_l2(_one(gen, u0), _one(burgers, u0)) <= TOL
    kdv = ex.stepper.KortewegDeVries(1, 3.0, 64, 0.01)
    gen = g.GeneralConvectionStepper(1, 3.0, 64, 0.01, linear_coefficients=[0.0, 0.0, 0.0, -1.0, -0.01],
                                     convection_scale=-6.0)
    assert max_rel_l2(_one(gen, u0), _one(kdv, u0)) <= TOL
    u0, _ = ics(3, 1, 60.0, 128, 5, seed=63)
    ks = ex.stepper.KuramotoSivashinsky(1, 60.0, 128, 0.1)
    gen = g.GeneralGradientNormStepper(1, 60.0, 128, 0.1, linear_coefficients=[0.0, 0.0, -1.0, 0.0, -1.0],
                                       gradient_norm_scale=1.0)
    assert max_rel_l2(_one(gen, u0), _one(ks, u0)) <= TOL


def test_normalized_matches_regular(device):
    L, N, dt, nu = 3.0, 64, 0.1, 0.1
    u0, _ = ics(3, 1, L, N, 5, seed=64)
    coefs = [0.0, -0.5, 0.01]
    reg = g.GeneralLinearStepper(1, L, N, dt, linear_coefficients=coefs)
    nor = g.NormalizedLinearStepper(1, N, normalized_linear_coefficients=g.normalize_coefficients(
        coefs, domain_extent=L, dt=dt))
    assert max_rel_l2(_one(nor, u0), _one(reg, u0)) <= TOL
    burgers = ex.stepper.Burgers(1, L, N, dt, diffusivity=nu)
    nb = g.NormalizedConvectionStepper(
        1, N, normalized_linear_coefficients=g.normalize_coefficients([0.0, 0.0, nu], domain_extent=L, dt=dt),
        normalized_convection_scale=g.normalize_convection_scale(1.0, domain_extent=L, dt=dt))
    assert max_rel_l2(_one(nb, u0), _one(burgers, u0)) <= TOL


def test_difficulty_steppers_run(device):
    import torch

    for st in (g.DifficultyConvectionStepper(num_points=64), g.DifficultyGradientNormStepper(num_points=64),
               g.DifficultyPolynomialStepper(num_points=64), g.DifficultyNonlinearStepper(num_points=64),
               g.DifficultyLinearStepper(num_points=64), g.DifficultyLinearStepperSimple(num_points=64)):
        u0, _ = ics(2, 1, 1.0, 64, 3, seed=65)
        out = st.rollout_batch(torch.tensor(u0, device=device), 5).cpu().numpy()
        assert np.isfinite(out).all(), type(st).__name__
