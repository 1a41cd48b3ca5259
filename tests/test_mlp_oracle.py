"""Pins of the surrogate network's CPU oracle (oracle/mlp_oracle.py).

The network has no counterpart in the reference, so the oracle is pinned by
what the mathematics fixes: central finite differences of the FP64 loss for
every gradient, hand-computed values for a one-unit network, torch's own
float32 -> bfloat16 rounding for the BF16 emulation, and Adam's first step,
which moves every weight by lr * sign(g) (up to eps).
"""
import math

import numpy as np
import pytest
import torch

import mlp_oracle as mo


def _net(dims, seed):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal((dims[i + 1], dims[i])) * 0.7, rng.standard_normal(dims[i + 1]) * 0.3)
            for i in range(len(dims) - 1)]


@pytest.mark.parametrize("dims", [[3, 5, 4, 1], [4, 1], [2, 6, 1]])
def test_gradients_match_finite_differences(dims):
    params = _net(dims, 1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((7, dims[0]))
    t = rng.standard_normal(7)
    loss, grads, _ = mo.loss_and_grads(params, x, t, round_bf16=False)
    h = 1e-6
    for li, (W, b) in enumerate(params):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            num = np.zeros_like(arr)
            it = np.nditer(arr, flags=["multi_index"])
            for _ in it:
                i = it.multi_index
                old = arr[i]
                arr[i] = old + h
                lp = mo.loss_and_grads(params, x, t, round_bf16=False)[0]
                arr[i] = old - h
                lm = mo.loss_and_grads(params, x, t, round_bf16=False)[0]
                arr[i] = old
                num[i] = (lp - lm) / (2 * h)
            assert np.allclose(g, num, rtol=1e-6, atol=1e-8), f"layer {li}: {np.max(np.abs(g - num))}"


def test_one_unit_network_by_hand():
    # y = w2 * silu(w1 x + b1) + b2 with x = 0.5, w1 = 2, b1 = -0.25, w2 = -1.5, b2 = 0.1
    params = [(np.array([[2.0]]), np.array([-0.25])), (np.array([[-1.5]]), np.array([0.1]))]
    y, _ = mo.forward(params, np.array([[0.5]]), round_bf16=False)
    pre = 0.75
    expect = -1.5 * pre / (1 + math.exp(-pre)) + 0.1
    assert abs(y[0] - expect) < 1e-15
    # zero weights: output is the head bias
    z = [(np.zeros((4, 3)), np.ones(4)), (np.zeros((1, 4)), np.array([0.7]))]
    y, _ = mo.forward(z, np.ones((5, 3)), round_bf16=True)
    assert np.all(y == 0.7)


def test_silu_grad_closed_form():
    x = np.linspace(-8, 8, 101)
    s = 1 / (1 + np.exp(-x))
    assert np.allclose(mo.silu_grad(x), s + x * s * (1 - s), rtol=1e-14, atol=1e-16)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.standard_normal(100000).astype(np.float32) * 10 ** rng.uniform(-5, 5, 100000),
                        np.array([1.0, -0.0, 0.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 65504.0], dtype=np.float32)])
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(a), want)


def test_adam_first_step_is_sign_step():
    rng = np.random.default_rng(4)
    w = rng.standard_normal(50)
    g = rng.standard_normal(50)
    w0 = w.copy()
    m = np.zeros(50)
    v = np.zeros(50)
    mo.adam(w, g, m, v, lr=1e-3, b1=0.9, b2=0.999, eps=0.0, step=1)
    assert np.allclose(w - w0, -1e-3 * np.sign(g), rtol=1e-12)
