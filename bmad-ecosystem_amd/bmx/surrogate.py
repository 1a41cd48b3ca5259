"""Learned dynamic-aperture surrogate: an MLP that maps (sextupole knob vector,
x0, y0, pz0) to log survival turns, trained on survival data from the batched
FP64 tracker and used to screen lattice variants, with every initial condition
near its predicted DA boundary re-tracked in FP64.

No counterpart in the reference (its DA search is the per-ray bisection of
bmad/modules/dynamic_aperture_mod.f90:27-398).  The network runs on libbmx's
own kernels (bmx_mlp.hip): BF16 matrix-core GEMMs with FP32 accumulation and
fused bias/SiLU epilogues for forward and backward, a column reduction for
bias gradients, an MSE-gradient kernel and an Adam step.  PyTorch only holds
device memory.
"""
import ctypes as C

import numpy as np

from . import abi

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class MLP:
    """Fully connected network dims[0] -> dims[1] -> ... -> dims[-1] = 1 with
    SiLU after every hidden layer and a linear head.

    Parameters live in one flat FP32 master buffer (per layer W [out, in] then
    b [out]) with a BF16 copy that the GEMMs read; gradients and Adam moments
    use the same layout, so one Adam launch updates everything.
    """

    def __init__(self, ctx, dims, params=None, seed=0):
        if dims[-1] != 1:
            raise ValueError("the head must have one output")
        self.ctx = ctx
        self.lib = ctx.lib
        self.dims = list(dims)
        self.dev = torch.device(f"cuda:{ctx.device}")
        self.layout = []
        off = 0
        for i in range(len(dims) - 1):
            n_in, n_out = dims[i], dims[i + 1]
            self.layout.append((off, off + n_out * n_in, n_in, n_out))
            off += n_out * n_in + n_out
        self.n_params = off
        z = lambda dt: torch.zeros(off, dtype=dt, device=self.dev)
        self.w, self.wb, self.g = z(torch.float32), z(torch.bfloat16), z(torch.float32)
        self.m, self.v = z(torch.float32), z(torch.float32)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self.step_count = 0
        self._bufs = {}
        if params is None:
            from_seed = _init_params(dims, seed)
            self.set_params(from_seed)
        else:
            self.set_params(params)

    # -- parameters -----------------------------------------------------------------
    def set_params(self, params):
        flat = np.zeros(self.n_params, dtype=np.float32)
        for (wo, bo, n_in, n_out), (W, b) in zip(self.layout, params):
            flat[wo:bo] = np.asarray(W, dtype=np.float32).reshape(-1)
            flat[bo:bo + n_out] = np.asarray(b, dtype=np.float32)
        torch.cuda.synchronize(self.dev)
        self.w.copy_(torch.from_numpy(flat))
        torch.cuda.synchronize(self.dev)
        self._check(self.lib.bmx_f32_to_bf16(self.ctx.h, self.n_params, _p(self.w), _p(self.wb)), "bmx_f32_to_bf16")
        self.sync()
        self.m.zero_()
        self.v.zero_()
        torch.cuda.synchronize(self.dev)
        self.step_count = 0

    def _unflat(self, t):
        a = t.detach().cpu().numpy().astype(np.float64)
        return [(a[wo:bo].reshape(n_out, n_in), a[bo:bo + n_out].copy()) for wo, bo, n_in, n_out in self.layout]

    def params(self):
        self.sync()
        return self._unflat(self.w)

    def grads(self):
        self.sync()
        return self._unflat(self.g)

    # -- plumbing --------------------------------------------------------------------
    def sync(self):
        self._check(self.lib.bmx_ctx_sync(self.ctx.h), "bmx_ctx_sync")

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {self.lib.bmx_last_error(self.ctx.h).decode()}")

    def _buffers(self, n):
        b = self._bufs.get(n)
        if b is None:
            f32 = lambda *s: torch.empty(*s, dtype=torch.float32, device=self.dev)
            b16 = lambda *s: torch.empty(*s, dtype=torch.bfloat16, device=self.dev)
            hid = self.dims[1:-1]
            wmax = max(hid) if hid else 1
            b = dict(
                xb=b16(n, self.dims[0]),
                pre=[f32(n, w) for w in hid],
                h=[b16(n, w) for w in hid],
                y=f32(n),
                d=f32(n), db=b16(n),
                dpre=[f32(n, wmax), f32(n, wmax)],
                dpre_b=[b16(n, wmax), b16(n, wmax)],
            )
            self._bufs = {n: b}  # keep one batch size resident
        return b

    def _gemm(self, ta, tb, M, N, K, A, lda, B, ldb, Cf, Cb, ldc, bias, epi, aux):
        rc = self.lib.bmx_gemm_bf16(self.ctx.h, ta, tb, M, N, K, A, lda, B, ldb, Cf, Cb, ldc, bias, epi, aux)
        self._check(rc, "bmx_gemm_bf16")

    def _wptr(self, layer, bf16=True):
        wo, bo, n_in, n_out = self.layout[layer]
        base = self.wb if bf16 else self.w
        return C.c_void_p(base.data_ptr() + wo * base.element_size())

    def _bptr(self, layer, t=None):
        wo, bo, n_in, n_out = self.layout[layer]
        t = self.w if t is None else t
        return C.c_void_p(t.data_ptr() + bo * t.element_size())

    def _gwptr(self, layer):
        wo, bo, n_in, n_out = self.layout[layer]
        return C.c_void_p(self.g.data_ptr() + wo * 4)

    # -- forward / backward ----------------------------------------------------------
    def forward(self, x):
        """x: FP32 device tensor [n, dims[0]] (torch work on it must be
        complete).  Returns the FP32 device tensor y [n] (valid after sync())."""
        n = x.shape[0]
        if x.dtype != torch.float32 or not x.is_contiguous() or x.shape[1] != self.dims[0]:
            raise ValueError("x must be a contiguous float32 [n, n_in] tensor")
        b = self._buffers(n)
        self._check(self.lib.bmx_f32_to_bf16(self.ctx.h, n * self.dims[0], _p(x), _p(b["xb"])), "bmx_f32_to_bf16")
        h = b["xb"]
        L = len(self.layout) - 1
        for l in range(L):
            wo, bo, n_in, n_out = self.layout[l]
            self._gemm(0, 1, n, n_out, n_in, _p(h), n_in, self._wptr(l), n_in, _p(b["pre"][l]), _p(b["h"][l]), n_out,
                       self._bptr(l), 2, None)
            h = b["h"][l]
        wo, bo, n_in, n_out = self.layout[L]
        self._gemm(0, 1, n, 1, n_in, _p(h), n_in, self._wptr(L), n_in, _p(b["y"]), None, 1, self._bptr(L), 1, None)
        return b["y"]

    def backward(self, target):
        """Gradients of the MSE loss of the last forward() against target
        (FP32 device tensor [n]) into self.g.  Returns the loss tensor."""
        n = target.shape[0]
        b = self._bufs[n]
        L = len(self.layout) - 1
        self._check(self.lib.bmx_mse_grad(self.ctx.h, n, _p(b["y"]), _p(target), _p(b["d"]), _p(b["db"]),
                                          _p(self.loss)), "bmx_mse_grad")
        # head: dW = d^T h_L, db = sum d, dh_L = d W_L -> dpre_{L-1}
        wo, bo, n_in, n_out = self.layout[L]
        h_in = b["h"][L - 1] if L > 0 else b["xb"]
        self._gemm(1, 0, 1, n_in, n, _p(b["db"]), 1, _p(h_in), n_in, self._gwptr(L), None, n_in, None, 0, None)
        self._check(self.lib.bmx_colsum_f32(self.ctx.h, n, 1, _p(b["d"]), 1, self._bptr(L, self.g)), "bmx_colsum_f32")
        if L == 0:
            return self.loss
        cur = 0
        self._gemm(0, 0, n, n_in, 1, _p(b["db"]), 1, self._wptr(L), n_in, _p(b["dpre"][cur]), _p(b["dpre_b"][cur]),
                   n_in, None, 3, _p(b["pre"][L - 1]))
        for l in range(L - 1, -1, -1):
            wo, bo, n_in, n_out = self.layout[l]
            h_in = b["h"][l - 1] if l > 0 else b["xb"]
            dp, dpb = b["dpre"][cur], b["dpre_b"][cur]
            self._gemm(1, 0, n_out, n_in, n, _p(dpb), n_out, _p(h_in), n_in, self._gwptr(l), None, n_in, None, 0,
                       None)
            self._check(self.lib.bmx_colsum_f32(self.ctx.h, n, n_out, _p(dp), n_out, self._bptr(l, self.g)),
                        "bmx_colsum_f32")
            if l > 0:
                nxt = 1 - cur
                self._gemm(0, 0, n, n_in, n_out, _p(dpb), n_out, self._wptr(l), n_in, _p(b["dpre"][nxt]),
                           _p(b["dpre_b"][nxt]), n_in, None, 3, _p(b["pre"][l - 1]))
                cur = nxt
        return self.loss

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        rc = self.lib.bmx_adam_step(self.ctx.h, self.n_params, _p(self.w), _p(self.wb), _p(self.g), _p(self.m),
                                    _p(self.v), lr, beta1, beta2, eps, self.step_count)
        self._check(rc, "bmx_adam_step")

    def train_step(self, x, t, lr):
        self.forward(x)
        loss = self.backward(t)
        self.adam(lr)
        return loss


def _init_params(dims, seed):
    """He-uniform weights (last layer scaled by 0.1), zero biases."""
    rng = np.random.default_rng(seed)
    ps = []
    for i in range(len(dims) - 1):
        lim = np.sqrt(6.0 / dims[i])
        W = rng.uniform(-lim, lim, size=(dims[i + 1], dims[i]))
        if i == len(dims) - 2:
            W *= 0.1
        ps.append((W, np.zeros(dims[i + 1])))
    return ps
