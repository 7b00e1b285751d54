"""Pins of the corrector network's CPU restatement (oracle/cnn.py).

- The float64 form is checked against torch.nn.functional.conv1d with circular
  padding, GELU (tanh form) and torch autograd on the CPU.
- The bfloat16 rounding helper is checked against torch's bfloat16 cast.
- A planted mistake (a shift of the kernel window by one) is caught.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import cnn


def torch_net(params, x):
    p = torch.as_tensor(params, dtype=torch.float64)
    layers, o = [], 0
    for ws, bs in cnn.SHAPES:
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        layers.append((p[o:o + nw].reshape(ws), p[o + nw:o + nw + nb]))
        o += nw + nb
    a = x
    for i, (W, b) in enumerate(layers):
        a = F.conv1d(F.pad(a, (2, 2), mode="circular"), W, b)
        if i < len(layers) - 1:
            a = F.gelu(a, approximate="tanh")
    return x + a


def test_param_count():
    assert cnn.NUM_PARAMS == 15809
    assert cnn.init_params(0).shape == (15809,)


def test_forward_matches_torch():
    rng = np.random.default_rng(0)
    params = cnn.init_params(1, scale_last=1.0) + 0.01 * rng.standard_normal(cnn.NUM_PARAMS)
    x = rng.standard_normal((3, 1, 64))
    y, _ = cnn.forward(params, x, bf16_model=False)
    yt = torch_net(params, torch.tensor(x)).numpy()
    np.testing.assert_allclose(y, yt, rtol=1e-12, atol=1e-12)


def test_backward_matches_torch_autograd():
    rng = np.random.default_rng(1)
    params = cnn.init_params(2, scale_last=1.0) + 0.01 * rng.standard_normal(cnn.NUM_PARAMS)
    x = rng.standard_normal((2, 1, 32))
    dy = rng.standard_normal((2, 1, 32))
    g, dx = cnn.backward(params, x, dy, bf16_model=False)
    pt = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    xt = torch.tensor(x, requires_grad=True)
    (torch_net(pt, xt) * torch.tensor(dy)).sum().backward()
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(dx, xt.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    v = (rng.standard_normal(10000) * 10.0 ** rng.uniform(-5, 5, 10000)).astype(np.float32)
    v[:4] = [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0]  # ties to even
    ref = torch.tensor(v).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(cnn.bf16(v), ref)


def test_bf16_model_close_to_float64():
    rng = np.random.default_rng(4)
    params = cnn.init_params(5, scale_last=1.0)
    x = rng.standard_normal((4, 1, 64))
    y64, _ = cnn.forward(params, x, bf16_model=False)
    y16, _ = cnn.forward(params, x, bf16_model=True)
    f64, f16 = y64 - x, y16 - x
    err = np.linalg.norm(f16 - f64) / np.linalg.norm(f64)
    assert 1e-4 < err < 3e-2  # bfloat16-level, not zero


def test_planted_window_shift_is_caught():
    rng = np.random.default_rng(6)
    params = cnn.init_params(7, scale_last=1.0)
    x = rng.standard_normal((2, 1, 32))
    good = torch_net(params, torch.tensor(x)).numpy()
    orig = cnn.conv
    try:
        cnn.conv = lambda a, W, b: np.roll(orig(a, W, b), 1, axis=-1)
        bad, _ = cnn.forward(params, x, bf16_model=False)
    finally:
        cnn.conv = orig
    assert np.abs(bad - good).max() > 1e-3
