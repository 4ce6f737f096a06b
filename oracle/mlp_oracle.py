"""CPU float64 restatement of the learned-closure network (include/helix_mlp.h).

TEST INFRASTRUCTURE ONLY: the checker for the HIP forward/backward kernels.  Same architecture and
parameter layout [W1 b1 W2 b2 W3 b3] (W stored [out][in]), tanh hidden layers, MSE loss averaged over
every output.  With bf16=True it restates the device's arithmetic: every GEMM runs on BF16 matrix cores,
so both operands of each product -- weights, inputs, activations and back-propagated errors -- are rounded
to BF16 (round to nearest even, from their FP32 values) and the products are accumulated exactly (the
device accumulates in FP32); biases, tanh and the error terms stay unrounded.  With bf16=False everything
is float64, and the backward pass is pinned by central finite differences (tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import numpy as np


def bf16_round(w: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even bfloat16, returned as float64 values."""
    f = np.asarray(w, np.float32)
    bits = f.view(np.uint32).astype(np.uint64)
    rounded = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, din: int, h1: int, h2: int, dout: int):
    p = np.asarray(params, np.float64)
    o = 0
    W1 = p[o:o + din * h1].reshape(h1, din); o += din * h1
    b1 = p[o:o + h1]; o += h1
    W2 = p[o:o + h1 * h2].reshape(h2, h1); o += h1 * h2
    b2 = p[o:o + h2]; o += h2
    W3 = p[o:o + h2 * dout].reshape(dout, h2); o += h2 * dout
    b3 = p[o:o + dout]; o += dout
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def _operand(bf16: bool):
    return bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))


def forward(params, x, dims, bf16=True):
    W1, b1, W2, b2, W3, b3 = unpack(params, *dims)
    r = _operand(bf16)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    a1 = np.tanh(r(x) @ W1.T + b1)
    a2 = np.tanh(r(a1) @ W2.T + b2)
    return r(a2) @ W3.T + b3, (a1, a2, W1, W2, W3)


def loss_and_grad(params, x, t, dims, bf16=True):
    """MSE loss (mean over n * out) and its gradient with respect to the packed parameters."""
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    r = _operand(bf16)
    y, (a1, a2, W1, W2, W3) = forward(params, x, dims, bf16)
    n, dout = y.shape
    res = y - t
    loss = float(np.sum(res * res) / (n * dout))
    d3 = 2.0 * res / (n * dout)
    gW3 = r(d3).T @ r(a2)
    gb3 = d3.sum(0)
    d2 = (r(d3) @ W3) * (1.0 - a2 * a2)
    gW2 = r(d2).T @ r(a1)
    gb2 = d2.sum(0)
    d1 = (r(d2) @ W2) * (1.0 - a1 * a1)
    gW1 = r(d1).T @ r(x)
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
