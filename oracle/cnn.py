"""CPU restatement of the learned corrector network (test infrastructure only).

The reference has no corrector. Its closest analogue is the periodic
convolutional emulator of
docs/examples/learning_burgers_autoregressive_neural_operator.ipynb (cells
4-39: a stack of circularly padded Conv1d layers). The network here is:
- 5 periodic Conv1d layers, kernel 5, widths 1 -> 32 -> 32 -> 32 -> 32 -> 1;
- GELU (tanh form) after layers 1-4;
- a residual output: y = x + f(x).

Arithmetic model of the GPU kernels, restated here:
- weights and the activations entering every convolution are rounded to
  bfloat16 (round to nearest even);
- products are accumulated in float32 (float64 here, which changes only the
  rounding of the sums), and biases are added in float32;
- GELU and its derivative are evaluated on the float32 pre-activation.

``bf16=False`` gives the plain float64 network. That is the form pinned
against torch.nn.functional.conv1d (circular padding) and autograd in
tests/test_oracle_cnn.py.

Parameter vector layout (float32, 15809 values), in order:
W1 (32,1,5), b1 (32), W2 (32,32,5), b2, W3, b3, W4, b4, W5 (1,32,5), b5 (1).
"""
from __future__ import annotations

import numpy as np

CH = 32
KW = 5
SHAPES = [((CH, 1, KW), (CH,)), ((CH, CH, KW), (CH,)), ((CH, CH, KW), (CH,)), ((CH, CH, KW), (CH,)),
          ((1, CH, KW), (1,))]
NUM_PARAMS = sum(int(np.prod(w)) + int(np.prod(b)) for w, b in SHAPES)
GELU_C = float(np.sqrt(2.0 / np.pi))
GELU_A = 0.044715


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned
    as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    bits = a.view(np.uint32).astype(np.uint64)
    rnd = ((bits >> 16) & 1) + 0x7FFF
    out = ((bits + rnd) & 0xFFFF0000).astype(np.uint32)
    return out.view(np.float32).reshape(a.shape)


def unpack(params):
    """Split the flat parameter vector into [(W, b), ...]."""
    p = np.asarray(params)
    out, o = [], 0
    for ws, bs in SHAPES:
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        out.append((p[o:o + nw].reshape(ws), p[o + nw:o + nw + nb].reshape(bs)))
        o += nw + nb
    return out


def init_params(seed=0, scale_last=0.1):
    """He-style initialisation; the last layer is scaled down so that the
    untrained network is a small correction."""
    rng = np.random.default_rng(seed)
    ps = []
    for i, (ws, bs) in enumerate(SHAPES):
        fan_in = ws[1] * ws[2]
        w = rng.standard_normal(ws) * np.sqrt(2.0 / fan_in)
        if i == len(SHAPES) - 1:
            w = w * scale_last
        ps += [w.ravel(), np.zeros(int(np.prod(bs)))]
    return np.concatenate(ps).astype(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def conv(a, W, b):
    """Periodic convolution (cross-correlation), 'same' size, kernel 5:
    out[s, o, p] = b[o] + sum_{i,k} W[o, i, k] a[s, i, (p + k - 2) mod N]."""
    N = a.shape[-1]
    out = np.zeros((a.shape[0], W.shape[0], N))
    for k in range(KW):
        out += np.einsum("oi,sip->sop", W[:, :, k], np.roll(a, -(k - 2), axis=-1))
    return out + b[None, :, None]


def conv_wgrad(a, d):
    """dW[o, i, k] = sum_{s,p} d[s, o, p] a[s, i, (p + k - 2) mod N]."""
    return np.stack([np.einsum("sop,sip->oi", d, np.roll(a, -(k - 2), axis=-1)) for k in range(KW)], axis=-1)


def conv_dgrad(d, W):
    """Transpose of conv in its input:
    da[s, i, q] = sum_{o,k} W[o, i, k] d[s, o, (q - k + 2) mod N]."""
    N = d.shape[-1]
    out = np.zeros((d.shape[0], W.shape[1], N))
    for k in range(KW):
        out += np.einsum("oi,sop->sip", W[:, :, k], np.roll(d, k - 2, axis=-1))
    return out


def forward(params, x, *, bf16_model=True):
    """y = x + f(x) for x of shape (S, 1, N). Returns (y, cache)."""
    r = bf16 if bf16_model else (lambda v: np.asarray(v, dtype=np.float64))
    layers = [(r(W).astype(np.float64), np.asarray(b, np.float64)) for W, b in unpack(params)]
    x = np.asarray(x, np.float64)
    a = [r(x).astype(np.float64)]
    z = []
    for li, (W, b) in enumerate(layers):
        zz = conv(a[-1], W, b)
        if li < len(layers) - 1:
            z.append(zz)
            a.append(r(gelu(zz)).astype(np.float64))
        else:
            out = zz
    return x + out, {"a": a, "z": z, "layers": layers}


def backward(params, x, dy, *, bf16_model=True):
    """Gradients of sum(y * dy): (d params (flat), dx). The activation
    gradients entering each transposed convolution are rounded to bfloat16
    in the bf16 model, as in the kernels."""
    r = bf16 if bf16_model else (lambda v: np.asarray(v, dtype=np.float64))
    _, c = forward(params, x, bf16_model=bf16_model)
    a, z, layers = c["a"], c["z"], c["layers"]
    dy = np.asarray(dy, np.float64)
    grads = [None] * len(layers)
    # last layer
    W5, _ = layers[-1]
    grads[-1] = (conv_wgrad(a[-1], dy), dy.sum(axis=(0, 2)))
    d = r(conv_dgrad(dy, W5) * gelu_grad(z[-1])).astype(np.float64)  # delta of layer 4
    for li in range(len(layers) - 2, -1, -1):
        W, _ = layers[li]
        grads[li] = (conv_wgrad(a[li], d), d.sum(axis=(0, 2)))
        if li > 0:
            d = r(conv_dgrad(d, W) * gelu_grad(z[li - 1])).astype(np.float64)
        else:
            dx = dy + conv_dgrad(d, W)
    flat = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])
    return flat, dx
