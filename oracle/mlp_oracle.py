"""CPU oracle of the dynamic-aperture surrogate network (test infrastructure:
only tests/, the smoke entry and bench.py's CPU leg use it; the product runs the
HIP kernels of bmx_mlp.hip through libbmx).

The network has no counterpart in the reference; this module states in FP64
numpy what the GPU computes, with the BF16 rounding points of the GPU path
made explicit so that the two agree to FP32-accumulation noise:

  h_0 = bf16(x)
  pre_l = h_{l-1} W_l^T + b_l,   h_l = bf16(silu(pre_l))      (hidden layers)
  y = h_L W_out^T + b_out                                     (linear head)
  loss = mean((y - t)^2)

Weights enter the products as bf16(W) (the bf16 copy of the FP32 master
weights); biases are FP32.  Backward follows the same rounding points:
the output gradient and every pre-activation gradient are rounded to BF16
before they enter a product, and bias gradients are column sums of the FP32
gradients.  With round_bf16=False every rounding is skipped, which is the form
the finite-difference checks in tests/test_mlp_oracle.py pin.
"""
import numpy as np


def bf16(a):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(x):
    return x / (1.0 + np.exp(-x))


def silu_grad(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def forward(params, x, round_bf16=True):
    """params: list of (W [out, in], b [out]) FP64 arrays, last one the head.
    Returns (y [n], cache)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    h = r(x)
    hs, pres = [h], []
    for W, b in params[:-1]:
        pre = h @ r(W).T + b
        if round_bf16:
            pre = pre.astype(np.float32).astype(np.float64)  # stored as FP32
        h = r(silu(pre))
        pres.append(pre)
        hs.append(h)
    W, b = params[-1]
    y = (h @ r(W).T + b)[:, 0]
    return y, (hs, pres)


def loss_and_grads(params, x, t, round_bf16=True):
    """MSE loss and its gradients with respect to every (W, b)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, (hs, pres) = forward(params, x, round_bf16)
    n = len(t)
    e = y - t
    loss = float(np.sum(e * e) / n)
    d = (2.0 * e / n)[:, None]  # [n, 1]
    grads = [None] * len(params)
    L = len(params) - 1
    db = r(d)
    grads[L] = (db.T @ hs[L], d.sum(axis=0))
    dpre_b = db
    for l in range(L - 1, -1, -1):
        W_next = params[l + 1][0]
        dh = dpre_b @ r(W_next)
        dpre = dh * silu_grad(pres[l])
        dpre_b = r(dpre)
        grads[l] = (dpre_b.T @ hs[l], dpre.sum(axis=0))
    return loss, grads, y


def adam(w, g, m, v, lr, b1, b2, eps, step):
    """One Adam step (Kingma & Ba, ICLR 2015, Algorithm 1) in place; returns w."""
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    w[:] = w - lr * mh / (np.sqrt(vh) + eps)
    return w


def init_params(dims, seed):
    """He-uniform initial weights, zero biases (the GPU network starts from
    the same arrays)."""
    rng = np.random.default_rng(seed)
    ps = []
    for i in range(len(dims) - 1):
        lim = np.sqrt(6.0 / dims[i])
        W = rng.uniform(-lim, lim, size=(dims[i + 1], dims[i]))
        if i == len(dims) - 2:
            W *= 0.1
        ps.append((W, np.zeros(dims[i + 1])))
    return ps
