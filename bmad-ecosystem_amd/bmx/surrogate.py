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
        if x.dtype not in (torch.float32, torch.bfloat16) or not x.is_contiguous() or x.shape[1] != self.dims[0]:
            raise ValueError("x must be a contiguous float32 or bfloat16 [n, n_in] tensor")
        b = self._buffers(n)
        if x.dtype == torch.bfloat16:
            h = x
            b["xb_ext"] = x
        else:
            self._check(self.lib.bmx_f32_to_bf16(self.ctx.h, n * self.dims[0], _p(x), _p(b["xb"])),
                        "bmx_f32_to_bf16")
            h = b["xb"]
            b["xb_ext"] = b["xb"]
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
        h_in = b["h"][L - 1] if L > 0 else b["xb_ext"]
        self._gemm(1, 0, 1, n_in, n, _p(b["db"]), 1, _p(h_in), n_in, self._gwptr(L), None, n_in, None, 0, None)
        self._check(self.lib.bmx_colsum_f32(self.ctx.h, n, 1, _p(b["d"]), 1, self._bptr(L, self.g)), "bmx_colsum_f32")
        if L == 0:
            return self.loss
        cur = 0
        self._gemm(0, 0, n, n_in, 1, _p(b["db"]), 1, self._wptr(L), n_in, _p(b["dpre"][cur]), _p(b["dpre_b"][cur]),
                   n_in, None, 3, _p(b["pre"][L - 1]))
        for l in range(L - 1, -1, -1):
            wo, bo, n_in, n_out = self.layout[l]
            h_in = b["h"][l - 1] if l > 0 else b["xb_ext"]
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

    def fused_ok(self):
        """True when the one-kernel forward pass takes this network
        (bmx_mlp_forward: dims[0] <= 64, hidden layers 256 wide)."""
        d = self.dims
        return len(d) >= 3 and 1 <= d[0] <= 64 and all(w == 256 for w in d[1:-1])

    def _dims_c(self):
        return (C.c_int32 * len(self.dims))(*self.dims)

    def forward_fused(self, x, y=None):
        """Inference-only forward pass in one kernel (bmx_mlp_forward), equal
        bit for bit to forward(): x BF16 device tensor [n, dims[0]]; returns
        the FP32 device tensor y [n] (valid after sync())."""
        n = x.shape[0]
        if x.dtype != torch.bfloat16 or not x.is_contiguous() or x.shape[1] != self.dims[0]:
            raise ValueError("x must be a contiguous bfloat16 [n, n_in] tensor")
        if y is None:
            y = torch.empty(n, dtype=torch.float32, device=self.dev)
        rc = self.lib.bmx_mlp_forward(self.ctx.h, len(self.dims), self._dims_c(), _p(self.w), _p(self.wb), _p(x), n,
                                      _p(y))
        self._check(rc, "bmx_mlp_forward")
        return y

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


class Net:
    """A surrogate network held by the library (bmx_surrogate_*): the same
    network and kernels as MLP, with its training loop, inference and DA
    screening in C++ (each training step replayed as a hipGraph), so that a
    caller needs no Python.  Interface as far as the DA pipeline uses MLP:
    dims, fused_ok(), params(), and train() / predict_boundaries() take it."""

    def __init__(self, ctx, dims, params=None, seed=0):
        self.ctx = ctx
        self.lib = ctx.lib
        self.dims = list(dims)
        self.dev = torch.device(f"cuda:{ctx.device}")
        flat = _flatten(dims, _init_params(dims, seed) if params is None else params)
        n = C.c_int64(0)
        h = C.c_void_p()
        self.h = None
        rc = self.lib.bmx_surrogate_create(ctx.h, len(self.dims), self._dims_c(),
                                           flat.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n), C.byref(h))
        self._check(rc, "bmx_surrogate_create")
        self.h = h
        self.n_params = n.value
        ctx._nets.add(self)

    def _dims_c(self):
        return (C.c_int32 * len(self.dims))(*self.dims)

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {self.lib.bmx_last_error(self.ctx.h).decode()}")

    def fused_ok(self):
        d = self.dims
        return len(d) >= 3 and 1 <= d[0] <= 64 and all(w == 256 for w in d[1:-1])

    def params(self):
        a = np.zeros(self.n_params, dtype=np.float32)
        self._check(self.lib.bmx_surrogate_get_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float))),
                    "bmx_surrogate_get_params")
        out, off = [], 0
        for i in range(len(self.dims) - 1):
            n_in, n_out = self.dims[i], self.dims[i + 1]
            W = a[off:off + n_in * n_out].astype(np.float64).reshape(n_out, n_in)
            off += n_in * n_out
            out.append((W, a[off:off + n_out].astype(np.float64)))
            off += n_out
        return out

    def set_params(self, params):
        flat = _flatten(self.dims, params)
        self._check(self.lib.bmx_surrogate_set_params(self.h, flat.ctypes.data_as(C.POINTER(C.c_float))),
                    "bmx_surrogate_set_params")

    def train(self, X, Y, epochs=20, batch=4096, lr=2e-3, lr_final=2e-4, seed=0, eager=False):
        """bmx_surrogate_train on (X, Y) (numpy float32 or device tensors);
        returns the per-epoch mean sampled losses."""
        Xd = X if torch.is_tensor(X) else torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32))
        Yd = Y if torch.is_tensor(Y) else torch.from_numpy(np.ascontiguousarray(Y, dtype=np.float32))
        Xd = Xd.to(self.dev, torch.float32).contiguous()
        Yd = Yd.to(self.dev, torch.float32).contiguous()
        torch.cuda.synchronize(self.dev)
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        rc = self.lib.bmx_surrogate_train(self.h, _p(Xd), _p(Yd), Xd.shape[0], int(epochs), int(batch), float(lr),
                                          float(lr_final), int(seed) & ((1 << 64) - 1), 1 if eager else 0,
                                          losses.ctypes.data_as(C.POINTER(C.c_float)))
        self._check(rc, "bmx_surrogate_train")
        return [float(v) for v in losses[:epochs]]

    def infer(self, xb):
        """Network output for BF16 device rows xb [n, dims[0]] -> FP32 device tensor [n]."""
        y = torch.empty(xb.shape[0], dtype=torch.float32, device=self.dev)
        torch.cuda.synchronize(self.dev)
        self._check(self.lib.bmx_surrogate_infer(self.h, _p(xb), xb.shape[0], _p(y)), "bmx_surrogate_infer")
        self._check(self.lib.bmx_ctx_sync(self.ctx.h), "bmx_ctx_sync")
        return y

    def close(self):
        if self.h:
            self.lib.bmx_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _flatten(dims, params):
    """Per layer W [out, in] then b [out], FP32."""
    parts = []
    for (W, b), n_in, n_out in zip(params, dims[:-1], dims[1:]):
        parts.append(np.asarray(W, dtype=np.float32).reshape(n_out, n_in).reshape(-1))
        parts.append(np.asarray(b, dtype=np.float32).reshape(n_out))
    return np.ascontiguousarray(np.concatenate(parts))


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


# ---------------------------------------------------------------------------
# Dynamic-aperture scan with the surrogate
# ---------------------------------------------------------------------------

N_KNOBS = 32
KNOB_RANGE = 0.2


class ScanSpec:
    """Polar grid of a DA scan (the bmx_track_grid layout): n_angle angles in
    [0, pi], n_amp amplitudes amp_max_x (k + 1) / n_amp, n_pz momentum offsets
    in [pz_min, pz_max], n_turns turns; the ring: n_cells C1 cells with n_rf
    rf cavities, and with error_seed the seeded errors of the batched-DA
    variants (quad k1 x (1 + 1e-3 g), sextupole k2 x (1 + 5e-3 g), 50 um
    quad and sextupole offsets) under the knobs."""

    def __init__(self, n_turns=256, n_angle=16, n_amp=32, n_pz=3, amp_max_x=0.02, amp_ratio_y=0.5, pz_min=-0.01,
                 pz_max=0.01, n_cells=50, n_rf=0, error_seed=None):
        self.n_turns, self.n_angle, self.n_amp, self.n_pz = n_turns, n_angle, n_amp, n_pz
        self.amp_max_x, self.amp_ratio_y, self.pz_min, self.pz_max = amp_max_x, amp_ratio_y, pz_min, pz_max
        self.n_cells, self.n_rf, self.error_seed = n_cells, n_rf, error_seed

    @property
    def n_ic(self):
        return self.n_angle * self.n_amp * self.n_pz

    def angles(self):
        return np.pi * np.arange(self.n_angle) / max(self.n_angle - 1, 1)

    def pzs(self):
        return np.linspace(self.pz_min, self.pz_max, self.n_pz) if self.n_pz > 1 else np.array([self.pz_min])

    def offsets(self):
        """[n_pz, n_angle, n_amp, 3] = (x, y, pz) of every grid IC."""
        a = self.amp_max_x * (np.arange(self.n_amp) + 1) / self.n_amp
        th = self.angles()
        pz = self.pzs()
        out = np.zeros((self.n_pz, self.n_angle, self.n_amp, 3))
        out[..., 0] = np.cos(th)[None, :, None] * a[None, None, :]
        out[..., 1] = self.amp_ratio_y * np.sin(th)[None, :, None] * a[None, None, :]
        out[..., 2] = pz[:, None, None]
        return out


def knob_lattice(knobs, spec, n_rf=None):
    """C1-cell ring of spec.n_cells cells (spec.n_rf cavities, the seeded
    errors of spec.error_seed) whose sextupole i has its k2 multiplied by
    (1 + knobs[i % 32])."""
    from . import abi
    from . import lattice as latmod
    ns = nq = 2 * spec.n_cells
    n_rf = spec.n_rf if n_rf is None else n_rf
    if spec.error_seed is None:
        eles = latmod.ring_c1(n_cells=spec.n_cells, n_rf=n_rf)
    else:
        rng = np.random.default_rng(spec.error_seed)
        k1e = 1e-3 * rng.standard_normal(nq)
        k2e = 5e-3 * rng.standard_normal(ns)
        offs = {}
        for i in range(nq):
            offs[("q", i)] = tuple(50e-6 * rng.standard_normal(2))
        for i in range(ns):
            offs[("s", i)] = tuple(50e-6 * rng.standard_normal(2))
        eles = latmod.ring_c1(n_cells=spec.n_cells, n_rf=n_rf, k1_err=k1e, k2_err=k2e, offsets=offs,
                              seed=spec.error_seed + 1)
    i = 0
    for e in eles:
        if e.key == abi.KEY_SEXTUPOLE:
            e.k2 = float(np.float64(e.k2) * (1.0 + np.float64(knobs[i % N_KNOBS])))
            i += 1
    return eles


def upload_variants(ctx, knob_sets, spec):
    """One lattice upload holding a knob_lattice per knob set: the nominal
    ring once, its sextupole k2 scaled per set by (1 + knob) in the library
    (bmx_lattice_upload_scaled: the same product knob_lattice forms)."""
    from . import abi, tracker
    knob_sets = np.asarray(knob_sets, dtype=np.float64)
    base = knob_lattice(np.zeros(N_KNOBS), spec)
    sext = np.array([i for i, e in enumerate(base) if e.key == abi.KEY_SEXTUPOLE], dtype=np.int32)
    fam = np.arange(len(sext)) % N_KNOBS
    factors = 1.0 + knob_sets[:, fam]
    if ctx is None or not hasattr(ctx, "h"):  # a stand-in context (tests): the records themselves
        return _upload_records(ctx, base, sext, factors)
    return tracker.Lattice.scaled(ctx, base, abi.ELECTRON, abi.ATTR_K2, sext, factors)


def _upload_records(ctx, base, sext, factors):
    """The scaled variants as one list of element records, uploaded through
    ctx.lattice (the records bmx_lattice_upload_scaled forms)."""
    import ctypes as Cc
    from . import abi
    from . import lattice as latmod
    arr0 = latmod.to_array(base)
    size = Cc.sizeof(abi.EleRec)
    n_ele = len(base)
    rec = np.frombuffer(arr0, dtype=np.uint8).reshape(n_ele, size)
    n_var = len(factors)
    big = np.ascontiguousarray(np.tile(rec, (n_var, 1)))
    k2_off = abi.EleRec.k2.offset
    k2 = big[:, k2_off:k2_off + 8].copy().view(np.float64).reshape(n_var, n_ele)
    k2[:, sext] = k2[:, sext] * factors
    big[:, k2_off:k2_off + 8] = k2.reshape(-1, 1).view(np.uint8)
    arr = (abi.EleRec * (n_var * n_ele)).from_buffer(big)
    offs = [v * n_ele for v in range(n_var + 1)]
    lat = ctx.lattice(arr, abi.ELECTRON, variant_offsets=offs)
    lat._records = big  # the ctypes array views this buffer
    return lat


def features(knob_sets, spec):
    """Network inputs [n_lat * n_ic, 35] (float32, host): knobs / 0.2 then the
    grid inputs of geometry(); the same rows bmx_da_features forms on the device."""
    knob_sets = np.asarray(knob_sets, dtype=np.float64)
    geo = geometry(spec)
    n_lat = len(knob_sets)
    X = np.empty((n_lat, spec.n_ic, N_KNOBS + 3), dtype=np.float32)
    X[:, :, :N_KNOBS] = (knob_sets / KNOB_RANGE)[:, None, :]
    X[:, :, N_KNOBS:] = geo[None]
    return X.reshape(-1, N_KNOBS + 3)


TARGET_GAMMA = 4.0


def target(survival, n_turns, gamma=TARGET_GAMMA):
    """Network target: log survival turns scaled to [0, 1] and raised to the
    power gamma, y = (log(1 + s) / log(1 + N))^gamma -- a monotone map of log
    survival that spreads the top of the range (full survival at 1, a
    particle lost after N/5 turns at 0.74^gamma), so the fit concentrates on
    the survive-every-turn boundary the scan needs."""
    y = np.log1p(np.asarray(survival, dtype=np.float64)) / np.log1p(n_turns)
    return (y ** gamma).astype(np.float32)


def survival_of(y, n_turns, gamma=TARGET_GAMMA):
    """Inverse of target(): predicted survival turns."""
    return np.expm1(np.clip(y, 0, None) ** (1.0 / gamma) * np.log1p(n_turns))


def plain_scan(ctx, knob_sets, spec, turns_per_block=0):
    """Full-grid FP64 tracking of every knob set in one batched launch.
    Returns (survival [n_lat, n_pz, n_angle, n_amp], area [n_lat], timing dict
    with total / upload (lattice build + upload) / track seconds)."""
    import time
    from . import tracker
    t0 = time.perf_counter()
    lat = upload_variants(ctx, knob_sets, spec)
    t1 = time.perf_counter()
    surv, _, _, st = tracker.track_grid(lat, spec.n_turns, spec.n_angle, spec.n_amp, spec.n_pz, spec.amp_max_x,
                                        spec.amp_ratio_y, spec.pz_min, spec.pz_max, turns_per_block=turns_per_block,
                                        return_stats=True)
    t2 = time.perf_counter()
    lat.close()
    area = np.array([slice_area(s[spec.n_pz // 2], spec) for s in surv])
    return surv, area, dict(total=t2 - t0, upload=t1 - t0, track=t2 - t1, pet=st["pet"])


def training_data(ctx, knob_sets, spec, n_ic, seed=0, chunk=512, gamma=TARGET_GAMMA):
    """Training set of the surrogate from FP64 tracking: for every knob set,
    n_ic grid initial conditions drawn at random (without replacement) from
    the spec's grid, tracked for spec.n_turns turns (bmx_track_ics, chunk
    lattices per launch).  Returns (X [n_lat * n_ic, 35] float32 features,
    Y [n_lat * n_ic] float32 targets, timing dict total / track seconds and
    the live particle-element-turns)."""
    import time
    from . import tracker
    knob_sets = np.asarray(knob_sets, dtype=np.float64)
    n_lat = len(knob_sets)
    rng = np.random.default_rng(seed)
    off = spec.offsets().reshape(-1, 3)
    geo = geometry(spec)
    X = np.empty((n_lat, n_ic, N_KNOBS + 3), dtype=np.float32)
    Y = np.empty((n_lat, n_ic), dtype=np.float32)
    t = dict(total=0.0, track=0.0, pet=0.0)
    for a in range(0, n_lat, chunk):
        b = min(n_lat, a + chunk)
        t0 = time.perf_counter()
        lat = upload_variants(ctx, knob_sets[a:b], spec)
        picks = [rng.choice(spec.n_ic, size=n_ic, replace=False) for _ in range(b - a)]
        per = []
        for idx in picks:
            c = np.zeros((6, n_ic))
            c[0], c[2], c[5] = off[idx, 0], off[idx, 1], off[idx, 2]
            per.append(c)
        t1 = time.perf_counter()
        surv, st = tracker.track_ics(lat, per, spec.n_turns, return_stats=True)
        t2 = time.perf_counter()
        lat.close()
        for j, (idx, sv) in enumerate(zip(picks, surv)):
            X[a + j, :, :N_KNOBS] = knob_sets[a + j] / KNOB_RANGE
            X[a + j, :, N_KNOBS:] = geo[idx]
            Y[a + j] = target(sv, spec.n_turns, gamma)
        t["total"] += t2 - t0
        t["track"] += t2 - t1
        t["pet"] += st["pet"]
    return X.reshape(-1, N_KNOBS + 3), Y.reshape(-1), t


def slice_area(surv_slice, spec):
    """DA area of one pz slice survival[n_angle, n_amp] (bmx_track_grid's definition)."""
    from . import da
    bx, by, _ = da.grid_boundary(surv_slice, spec.n_turns, spec.amp_max_x, spec.amp_ratio_y, spec.angles())
    return da.polygon_area(bx, by)


def area_from_runs(runs, spec):
    """Area from per-angle surviving-prefix lengths runs[n_angle]."""
    from . import da
    a = spec.amp_max_x * np.asarray(runs) / spec.n_amp
    th = spec.angles()
    return da.polygon_area(a * np.cos(th), a * spec.amp_ratio_y * np.sin(th))


def train(net, X, Y, epochs=20, batch=4096, lr=2e-3, seed=0, lr_final=2e-4, loss_every=16):
    """Minibatch Adam on (X, Y) (numpy float32) with a cosine-decayed learning
    rate; every epoch is one shuffled pass.  Each epoch gathers its shuffled
    copy of the data on the device once; its steps then run back to back on
    the network's stream, which the host joins every `loss_every` steps (and
    at the end of the epoch) to read the loss.  Returns the per-epoch means
    of those sampled losses."""
    if isinstance(net, Net):
        return net.train(X, Y, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, seed=seed)
    dev = net.dev
    Xd = torch.from_numpy(np.ascontiguousarray(X)).to(dev)
    Yd = torch.from_numpy(np.ascontiguousarray(Y)).to(dev)
    g = torch.Generator(device="cpu").manual_seed(seed)
    n = Xd.shape[0]
    nb = n // batch
    steps = epochs * max(1, nb)
    step = 0
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(dev)
        Xe = Xd.index_select(0, perm)
        Ye = Yd.index_select(0, perm)
        torch.cuda.synchronize(dev)  # the gathered epoch is complete before the network's stream reads it
        acc, cnt = 0.0, 0
        for k in range(nb):
            frac = step / max(steps - 1, 1)
            cur = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac))
            loss = net.train_step(Xe[k * batch:(k + 1) * batch], Ye[k * batch:(k + 1) * batch], float(cur))
            step += 1
            if k % loss_every == loss_every - 1 or k == nb - 1:
                net.sync()
                acc += float(loss.item())
                cnt += 1
        net.sync()  # every step has read Xe, Ye before they go
        del Xe, Ye
        losses.append(acc / max(cnt, 1))
    return losses


def predict(net, X, batch=1 << 16):
    """Network output for X (numpy float32 [n, 35]) -> numpy float32 [n]."""
    out = np.empty(len(X), dtype=np.float32)
    Xd = torch.from_numpy(np.ascontiguousarray(X)).to(net.dev)
    for b in range(0, len(X), batch):
        xb = Xd[b:b + batch].contiguous()
        torch.cuda.synchronize(net.dev)
        y = net.forward(xb)
        net.sync()
        out[b:b + len(xb)] = y.cpu().numpy()
    return out


def step_fit_boundary(pred, alive_level):
    """Boundary index of every ray from its predicted profile pred[..., n_amp]:
    the k minimising sum_{j<k} (p_j - alive_level)^2 + sum_{j>=k} (p_j - mean_{j>=k} p)^2,
    the least-squares fit of a step from the surviving level to a free lower
    level (uses the whole profile, so one mispredicted point does not move it).
    Returns k in 0..n_amp (n_amp: the whole ray survives)."""
    p = np.asarray(pred, dtype=np.float64)
    K = p.shape[-1]
    pre = np.concatenate([np.zeros(p.shape[:-1] + (1,)), np.cumsum((p - alive_level) ** 2, -1)], -1)  # k = 0..K
    rs = np.cumsum(p[..., ::-1], -1)[..., ::-1]
    rq = np.cumsum((p * p)[..., ::-1], -1)[..., ::-1]
    zero = np.zeros(p.shape[:-1] + (1,))
    suf_s = np.concatenate([rs, zero], -1)
    suf_q = np.concatenate([rq, zero], -1)
    cnt = (K - np.arange(K + 1)).astype(np.float64)
    suf = suf_q - np.where(cnt > 0, suf_s * suf_s / np.maximum(cnt, 1), 0.0)
    cost = pre + suf
    return K - np.argmin(cost[..., ::-1], -1)  # ties (a flat profile) resolve to the larger k


def geometry(spec):
    """[n_ic, 3] grid inputs: x0 / amp_max, y0 / (amp_max ratio_y), pz / max|pz|."""
    off = spec.offsets().reshape(-1, 3)
    pzs = max(abs(spec.pz_min), abs(spec.pz_max), 1e-30)
    return np.stack([off[:, 0] / spec.amp_max_x, off[:, 1] / (spec.amp_max_x * spec.amp_ratio_y), off[:, 2] / pzs], 1)


def predict_knobs(net, knob_sets, spec, lat_chunk=None):
    """Network output for every grid IC of every knob set
    [n_lat, n_pz, n_angle, n_amp] float32.  The inputs are formed on the
    device (bmx_da_features), one chunk of lattices at a time."""
    knob_sets = np.asarray(knob_sets)
    n_lat = len(knob_sets)
    if lat_chunk is None:
        lat_chunk = max(1, (1 << 21) // spec.n_ic)
    dev = net.dev
    geo = torch.from_numpy(geometry(spec).astype(np.float32)).to(dev)
    kd = torch.from_numpy(knob_sets.astype(np.float32)).to(dev)
    width = N_KNOBS + 3
    out = torch.empty(n_lat * spec.n_ic, dtype=torch.float32, device=dev)
    X = torch.empty(min(lat_chunk, n_lat) * spec.n_ic, width, dtype=torch.bfloat16, device=dev)
    torch.cuda.synchronize(dev)
    for a in range(0, n_lat, lat_chunk):
        b = min(n_lat, a + lat_chunk)
        rows = (b - a) * spec.n_ic
        xb = X[:rows]
        rc = net.lib.bmx_da_features(net.ctx.h, b - a, spec.n_ic, N_KNOBS, 3,
                                     C.c_void_p(kd.data_ptr() + a * N_KNOBS * 4), KNOB_RANGE, _p(geo), _p(xb))
        net._check(rc, "bmx_da_features")
        y = net.forward(xb)
        net.sync()
        out[a * spec.n_ic:b * spec.n_ic].copy_(y)
        torch.cuda.synchronize(dev)
    return out.cpu().numpy().reshape(n_lat, spec.n_pz, spec.n_angle, spec.n_amp)


def predict_boundaries(net, knob_sets, spec, lat_chunk=None, layered=False):
    """Predicted boundary index of every ray [n_lat, n_pz, n_angle] (int64,
    0..n_amp): the network output for every grid IC and its step fit
    (step_fit_boundary at the surviving level 1), all on the device; only the
    indices come back to the host.  A network the one-kernel forward pass
    takes runs through bmx_da_screen; otherwise (or with layered=True) the
    calls are bmx_da_features, the layer-by-layer forward and bmx_step_fit."""
    knob_sets = np.asarray(knob_sets)
    n_lat = len(knob_sets)
    if lat_chunk is None:
        lat_chunk = max(1, (1 << 21) // spec.n_ic)
    dev = net.dev
    geo = torch.from_numpy(geometry(spec).astype(np.float32)).to(dev)
    kd = torch.from_numpy(knob_sets.astype(np.float32)).to(dev)
    n_ray_lat = spec.n_pz * spec.n_angle
    kb = torch.empty(n_lat * n_ray_lat, dtype=torch.int32, device=dev)
    if isinstance(net, Net):
        torch.cuda.synchronize(dev)
        rc = net.lib.bmx_surrogate_screen(net.h, n_lat, N_KNOBS, _p(kd), KNOB_RANGE, spec.n_ic, 3, _p(geo),
                                          spec.n_amp, 1.0, _p(kb))
        net._check(rc, "bmx_surrogate_screen")
        net._check(net.lib.bmx_ctx_sync(net.ctx.h), "bmx_ctx_sync")
        return kb.cpu().numpy().astype(np.int64).reshape(n_lat, spec.n_pz, spec.n_angle)
    if net.fused_ok() and not layered:  # features, forward pass and step fit in the library
        torch.cuda.synchronize(dev)
        rc = net.lib.bmx_da_screen(net.ctx.h, len(net.dims), net._dims_c(), _p(net.w), _p(net.wb), n_lat, N_KNOBS,
                                   _p(kd), KNOB_RANGE, spec.n_ic, 3, _p(geo), spec.n_amp, 1.0, _p(kb))
        net._check(rc, "bmx_da_screen")
        net.sync()
        return kb.cpu().numpy().astype(np.int64).reshape(n_lat, spec.n_pz, spec.n_angle)
    X = torch.empty(min(lat_chunk, n_lat) * spec.n_ic, N_KNOBS + 3, dtype=torch.bfloat16, device=dev)
    torch.cuda.synchronize(dev)
    for a in range(0, n_lat, lat_chunk):
        b = min(n_lat, a + lat_chunk)
        xb = X[:(b - a) * spec.n_ic]
        rc = net.lib.bmx_da_features(net.ctx.h, b - a, spec.n_ic, N_KNOBS, 3,
                                     C.c_void_p(kd.data_ptr() + a * N_KNOBS * 4), KNOB_RANGE, _p(geo), _p(xb))
        net._check(rc, "bmx_da_features")
        y = net.forward(xb)
        rc = net.lib.bmx_step_fit(net.ctx.h, (b - a) * n_ray_lat, spec.n_amp, _p(y), 1.0,
                                  C.c_void_p(kb.data_ptr() + a * n_ray_lat * 4))
        net._check(rc, "bmx_step_fit")
        net.sync()
    return kb.cpu().numpy().astype(np.int64).reshape(n_lat, spec.n_pz, spec.n_angle)


def _resolve_rays(lat, n_lat, spec, lo, hi, refine, turns_per_block, max_rounds, up=0, down=0):
    """FP64 rounds of surrogate_scan for one chunk of n_lat lattices,
    uploaded as the variants of lat (closed on return): the window [lo, hi]
    of every ray first, then further rounds until every
    ray is resolved or max_rounds have run.  Bookkeeping is per ray:
    bnd = the first tracked loss (n_amp if none) and a_top = the highest
    tracked survivor below it (-1 if none); the ray is resolved when no
    untracked amplitude lies between them.  Later rounds: with up = down = 0,
    up to `refine` evenly spaced points of each open gap (refine = 0: the
    whole gap); with up, down > 0 (bracket mode), the `up` amplitudes above
    a_top of a ray with no loss yet, the `down` amplitudes below bnd of a ray
    with no survivor yet, and the whole gap of a ray holding both.
    Returns (bnd, a_top [n_lat, n_pz, n_angle], number of tracked ICs,
    rounds [(n_ics, seconds, pet)])."""
    import time
    from . import tracker
    N = spec.n_turns
    A = spec.n_amp
    n_ray_lat = spec.n_pz * spec.n_angle
    R = n_lat * n_ray_lat
    off = spec.offsets().reshape(n_ray_lat, A, 3)
    bnd = np.full(R, A, dtype=np.int64)
    a_top = np.full(R, -1, dtype=np.int64)
    tracked = np.zeros((R, A), dtype=bool)
    alive_r, alive_k = [], []  # every tracked survivor
    rounds = []
    # round 1: the band of every ray
    lo, hi = lo.reshape(-1), hi.reshape(-1)
    cnt = np.maximum(hi - lo + 1, 0)
    ray = np.repeat(np.arange(R), cnt)
    amp = lo[ray] + (np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    for rnd in range(max_rounds):
        keep = ~tracked[ray, amp]
        ray, amp = ray[keep], amp[keep]
        if len(ray) == 0:
            break
        # the launch takes the ICs lattice by lattice: order them by ray
        order = np.argsort(ray * A + amp, kind="stable")
        ray, amp = ray[order], amp[order]
        tracked[ray, amp] = True
        o = off[ray % n_ray_lat, amp]
        counts = np.bincount(ray // n_ray_lat, minlength=n_lat)
        c = np.zeros((6, len(ray)))
        c[0], c[2], c[5] = o[:, 0], o[:, 1], o[:, 2]
        per = np.split(c, np.cumsum(counts)[:-1], axis=1)
        tr0 = time.perf_counter()
        surv, st = tracker.track_ics(lat, per, N, turns_per_block, return_stats=True)
        rounds.append((len(ray), time.perf_counter() - tr0, st["pet"], st.get("kernel_ms", 0.0) / 1e3))
        surv = np.concatenate(surv)
        lost = surv < N
        np.minimum.at(bnd, ray[lost], amp[lost])
        alive_r.append(ray[~lost])
        alive_k.append(amp[~lost])
        ar, ak = np.concatenate(alive_r), np.concatenate(alive_k)
        below = ak < bnd[ar]
        a_top[:] = -1
        np.maximum.at(a_top, ar[below], ak[below])
        gap = bnd - a_top - 1
        open_r = np.nonzero(gap > 0)[0]
        if len(open_r) == 0 or rnd == max_rounds - 1:
            break
        g = gap[open_r]
        if up > 0 or down > 0:  # bracket mode
            no_loss, no_surv = bnd[open_r] == A, a_top[open_r] == -1
            rr, aa = [], []
            o = open_r[no_loss & ~no_surv]
            pts = a_top[o, None] + 1 + np.arange(up)[None, :]
            rr.append(np.broadcast_to(o[:, None], pts.shape)[pts < A])
            aa.append(pts[pts < A])
            o = open_r[no_surv & ~no_loss]
            pts = bnd[o, None] - 1 - np.arange(down)[None, :]
            rr.append(np.broadcast_to(o[:, None], pts.shape)[pts >= 0])
            aa.append(pts[pts >= 0])
            both = ~no_surv & ~no_loss
            o, gb = open_r[both], g[both]
            rr.append(np.repeat(o, gb))
            aa.append(np.repeat(a_top[o], gb) + 1 + (np.arange(gb.sum()) - np.repeat(np.cumsum(gb) - gb, gb)))
            ray, amp = np.concatenate(rr).astype(np.int64), np.concatenate(aa).astype(np.int64)
            continue
        if refine <= 0:  # the whole gap in one round
            ray = np.repeat(open_r, g)
            amp = a_top[ray] + 1 + (np.arange(g.sum()) - np.repeat(np.cumsum(g) - g, g))
            continue
        # up to `refine` evenly spaced points of each open gap (a_top, bnd)
        j = np.arange(1, refine + 1)
        pts = a_top[open_r, None] + np.ceil((g[:, None] + 1) * j / (refine + 1)).astype(np.int64)
        ok = (pts > a_top[open_r, None]) & (pts < bnd[open_r, None])
        ray = np.broadcast_to(open_r[:, None], pts.shape)[ok]
        amp = pts[ok]
        u = np.unique(ray * A + amp)
        ray, amp = u // A, u % A
    lat.close()
    shape = (n_lat, spec.n_pz, spec.n_angle)
    return bnd.reshape(shape), a_top.reshape(shape), int(tracked.sum()), rounds


def neighbour_estimate(bnd, a_top, kb, reach=3):
    """Boundary of every ray [..., n_angle]: bnd where the ray is resolved
    (a_top = bnd - 1), else the mean boundary of the resolved rays up to
    `reach` angles away on either side (weights 1/distance; the angles span
    0..pi, so the two ends are not neighbours), or the prediction kb where
    none is resolved, rounded and clipped into the ray's bracket
    [a_top + 1, bnd]."""
    res = a_top == bnd - 1
    b = np.where(res, bnd, 0).astype(np.float64)
    w = res.astype(np.float64)
    num = np.zeros_like(b)
    den = np.zeros_like(b)
    for d in range(1, reach + 1):
        num[..., d:] += b[..., :-d] / d  # neighbour d angles below
        den[..., d:] += w[..., :-d] / d
        num[..., :-d] += b[..., d:] / d  # neighbour d angles above
        den[..., :-d] += w[..., d:] / d
    est = np.where(den > 0, num / np.maximum(den, 1e-300), kb)
    est = np.clip(np.rint(est).astype(np.int64), a_top + 1, bnd)
    return np.where(res, bnd, est)


def surrogate_scan(ctx, net, knob_sets, spec, band=0.10, turns_per_block=0, max_rounds=64, refine=6, workers=1,
                   window=None, probe_up=4, probe_down=3):
    """Surrogate-screened DA scan.  Every ray (lattice, pz, angle) gets a
    predicted boundary index kb from the network (step_fit_boundary of its
    predicted profile).

    Exact mode (window None): the initial conditions whose amplitude lies
    within +-band of the predicted boundary amplitude are re-tracked in FP64
    (bmx_track_ics, one launch for all rays of a chunk).  Where the re-tracked
    band does not pin the first loss -- untracked amplitudes remain between
    the highest tracked survivor and the first tracked loss (or the grid
    edge) -- the next round tracks up to `refine` evenly spaced points of the
    gap (refine = 0: the whole gap in one round), until every ray is
    resolved: plain FP64 tracking is the fallback and the judge of every
    boundary.

    Bracket mode (window = (a, b)): round 1 tracks amplitudes kb-a .. kb+b of
    every ray; each later round probes probe_up amplitudes above the top
    survivor of a ray with no loss yet, probe_down below the first loss of a
    ray with no survivor yet, and the whole gap of a bracketed ray; after
    max_rounds a ray still open takes neighbour_estimate().  Every round is
    as long as its longest survivor (n_turns), so a bounded number of rounds
    bounds the scan time; the rays left open are a few per cent, scattered,
    and their neighbours in angle carry the boundary.

    Band mode (window None, max_rounds = 1): the one FP64 round of the band
    decides every ray, with the network standing for the untracked
    amplitudes: below the band they survive, so the surviving prefix ends at
    the first tracked loss; a band without a loss puts the boundary just
    above it (the lowest amplitude the FP64 round leaves open), capped at the
    grid's edge.

    workers > 1 splits the lattices into that many chunks resolved
    concurrently, each on its own context (HIP stream), so the
    latency-bound late rounds of one chunk overlap the others' work; their
    lattice uploads (host work) run while the network predicts.
    Returns (runs [n_lat, n_pz, n_angle], area [n_lat], timing dict
    (total / predict / track seconds, rounds, kb, n_open), n_tracked)."""
    import threading
    import time
    from . import tracker
    t0 = time.perf_counter()
    knob_sets = np.asarray(knob_sets)
    n_lat = len(knob_sets)
    workers = max(1, min(workers, n_lat))
    bounds = np.linspace(0, n_lat, workers + 1).astype(int)
    results = [None] * workers
    errors = []
    boxes = [{} for _ in range(workers)]
    ready = [threading.Event() for _ in range(workers)]
    ctxs = [ctx] if workers == 1 else [tracker.Context(ctx.device) for _ in range(workers)]
    amps = spec.amp_max_x * (np.arange(spec.n_amp) + 1) / spec.n_amp

    def window_of(kb):
        if window is None:  # band in index space: amplitudes within +-band of the boundary amplitude
            a_b = spec.amp_max_x * np.maximum(kb, 1) / spec.n_amp
            lo = np.searchsorted(amps, a_b * (1 - band), side="left")
            hi = np.searchsorted(amps, a_b * (1 + band), side="right") - 1
            lo = np.clip(np.minimum(lo, kb - 1), 0, spec.n_amp - 1)
            hi = np.clip(np.maximum(hi, kb), 0, spec.n_amp - 1)
            return dict(lo=lo, hi=hi, up=0, down=0)
        return dict(lo=np.clip(kb - window[0], 0, spec.n_amp - 1), hi=np.clip(kb + window[1], 0, spec.n_amp - 1),
                    up=max(1, probe_up), down=max(1, probe_down))

    timeline = [dict() for _ in range(workers)]

    def work(w):
        try:
            a, b = bounds[w], bounds[w + 1]
            tw0 = time.perf_counter()
            lat = upload_variants(ctxs[w], knob_sets[a:b], spec)
            tw1 = time.perf_counter()
            ready[w].wait()
            tw2 = time.perf_counter()
            bx = boxes[w]
            if "lo" not in bx:  # the prediction failed
                lat.close()
                return
            results[w] = _resolve_rays(lat, b - a, spec, bx["lo"], bx["hi"], refine, turns_per_block, max_rounds,
                                       bx["up"], bx["down"])
            tw3 = time.perf_counter()
            timeline[w].update(upload_end=tw1 - t0, ready=tw2 - t0, resolved=tw3 - t0)
        except Exception as e:  # re-raised in the caller
            errors.append(e)

    # the network predicts chunk by chunk, and each chunk's FP64 rounds start (on its own stream) while
    # the next chunk is predicted
    th = [threading.Thread(target=work, args=(w,)) for w in range(workers)] if workers > 1 else []
    for t in th:
        t.start()
    kbs = []
    try:
        for w in range(workers):
            kb_w = predict_boundaries(net, knob_sets[bounds[w]:bounds[w + 1]], spec)
            kbs.append(kb_w)
            boxes[w].update(window_of(kb_w))
            timeline[w]["predicted"] = time.perf_counter() - t0
            ready[w].set()
        t_pred = time.perf_counter()
    finally:
        for e in ready:
            e.set()
        if workers == 1:
            work(0)
        for t in th:
            t.join()
        if workers > 1:
            for c in ctxs:
                c.close()
    if errors:
        raise errors[0]
    kb = np.concatenate(kbs, 0)
    # surviving prefix: untracked amplitudes below the first tracked loss count as alive
    bnd = np.concatenate([r[0] for r in results], 0)
    a_top = np.concatenate([r[1] for r in results], 0)
    n_tracked = sum(r[2] for r in results)
    rounds = [x for r in results for x in r[3]]
    n_open = int(np.sum(a_top != bnd - 1))
    if window is None and max_rounds == 1:  # band mode
        hi = np.concatenate([boxes[w]["hi"].reshape(bnd[bounds[w]:bounds[w + 1]].shape) for w in range(workers)], 0)
        runs = np.where(bnd < spec.n_amp, bnd, np.minimum(hi + 1, spec.n_amp))
    else:
        runs = neighbour_estimate(bnd, a_top, kb)
    area = np.array([area_from_runs(runs[v, spec.n_pz // 2], spec) for v in range(n_lat)])
    t1 = time.perf_counter()
    timing = dict(total=t1 - t0, predict=t_pred - t0, track=t1 - t_pred, rounds=rounds, kb=kb, n_open=n_open,
                  timeline=timeline)
    return runs, area, timing, n_tracked


def inference_throughput(net, n=1 << 18, reps=10):
    """Dense MFMA throughput of the network (algorithmic BF16 flops / wall
    time between stream syncs) at batch n: the forward pass layer by layer
    (bmx_gemm_bf16 chain) and in one kernel (bmx_mlp_forward, when the shape
    allows), and training steps (forward, backward, Adam).  Measured on a
    copy of the network's parameters, which it leaves unchanged."""
    import time
    dims = net.dims
    m = MLP(net.ctx, dims, params=net.params())
    flop = 2 * sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
    x = torch.rand(n, dims[0], device=m.dev, dtype=torch.float32)
    xb = x.to(torch.bfloat16)
    t = torch.rand(n, device=m.dev, dtype=torch.float32)
    y = torch.empty(n, device=m.dev, dtype=torch.float32)
    torch.cuda.synchronize(m.dev)

    def timed(f):
        f()
        m.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        m.sync()
        return (time.perf_counter() - t0) / reps

    fwd = timed(lambda: m.forward(xb))
    fused = timed(lambda: m.forward_fused(xb, y)) if m.fused_ok() else None
    trn = timed(lambda: m.train_step(x, t, 0.0))
    out = {"forward_tflops": n * flop / fwd / 1e12, "train_tflops": 3 * n * flop / trn / 1e12,
           "forward_ms": fwd * 1e3, "train_step_ms": trn * 1e3, "batch": n}
    if fused is not None:
        out.update(forward_fused_tflops=n * flop / fused / 1e12, forward_fused_ms=fused * 1e3)
    return out


def area_errors(runs, plain_runs, spec):
    """DA-area errors of screened runs [n_lat, n_pz, n_angle] against plain
    ones: of the on-momentum slice (relative), of every pz slice (relative
    where the plain slice has area; absolute in units of the mean
    on-momentum area everywhere) and of the areas summed over the slices."""
    n_lat = len(runs)
    a_s = np.array([[area_from_runs(runs[v, p], spec) for p in range(spec.n_pz)] for v in range(n_lat)])
    a_p = np.array([[area_from_runs(plain_runs[v, p], spec) for p in range(spec.n_pz)] for v in range(n_lat)])
    mid = spec.n_pz // 2
    # relative errors where the plain area is not zero; lattices with no
    # on-momentum aperture at all are counted apart
    nz = a_p[:, mid] > 0
    err = np.abs(a_s[nz, mid] - a_p[nz, mid]) / a_p[nz, mid]
    if err.size == 0:
        err = np.zeros(1)
    has = a_p > 0
    err_all = np.abs(a_s - a_p)[has] / a_p[has]
    tot_s, tot_p = a_s.sum(1), a_p.sum(1)
    err_tot = np.abs(tot_s - tot_p) / np.maximum(tot_p, 1e-300)
    return {
        "da_area_err_mean": float(err.mean()), "da_area_err_max": float(err.max()),
        "frac_within_2pct": float(np.mean(err <= 0.02)),
        "da_area_err_all_slices_mean": float(err_all.mean()) if err_all.size else 0.0,
        "da_area_err_all_slices_frac_within_2pct": float(np.mean(err_all <= 0.02)) if err_all.size else 1.0,
        "da_area_err_all_slices_abs_max_rel_mean_area": float(np.abs(a_s - a_p).max() / np.mean(a_p[:, mid])),
        "da_area_err_summed_slices_mean": float(err_tot.mean()), "da_area_err_summed_slices_max": float(err_tot.max()),
        "rays_identical": float(np.mean(runs == plain_runs)),
        "area_plain_mean_mm2": float(a_p[:, mid].mean() * 1e6),
        "n_zero_area_plain": int(np.sum(~nz)), "zero_area_found": int(np.sum(~nz & (a_s[:, mid] == 0))),
    }


BAND_MODE = dict(window=None, max_rounds=1)
BRACKET_MODE = dict(window=(1, 0), probe_up=4, probe_down=3, max_rounds=2)


def compare_scans(ctx, spec, n_train=64, n_test=512, epochs=4, batch=16384, seed=20261205, band=0.10,
                  gamma=TARGET_GAMMA, workers=1, refine=6, window=None, probe_up=4, probe_down=3, max_rounds=64,
                  scan_chunk=512, train_ics=None, modes=None, lr=2e-3):
    """Surrogate-screened vs plain DA scans of n_test unseen knob settings.
    Trains a fresh 4 x 256 network on FP64 tracking of n_train settings:
    train_ics random grid ICs per setting (training_data; None: the whole
    grid, as plain scans in chunks of scan_chunk lattices).  Then scans the
    test settings plainly and with the network, once per entry of `modes`
    (dicts of surrogate_scan arguments; None: one mode from the window /
    max_rounds / probe arguments), on the same network and lattices.
    Returns a dict of timings, lattices/hour (lattice upload + tracking for
    plain; upload + inference + FP64 re-tracking for the surrogate),
    FP64-tracked IC counts and the DA-area errors (area_errors); with several
    modes their records are under "modes" and the first one's are also at
    the top level."""
    import sys
    import time
    rng = np.random.default_rng(seed)
    k_train = rng.uniform(-KNOB_RANGE, KNOB_RANGE, (n_train, N_KNOBS))
    k_test = rng.uniform(-KNOB_RANGE, KNOB_RANGE, (n_test, N_KNOBS))

    def say(msg):
        print(f"da comparison: {msg}", file=sys.stderr, flush=True)

    say(f"training data: {n_train} lattices x {train_ics or spec.n_ic} ICs, {spec.n_turns} turns")
    if train_ics is None:
        Y, t_gen = [], dict(total=0.0, track=0.0)
        for a in range(0, n_train, scan_chunk):
            s_tr, _, t = plain_scan(ctx, k_train[a:a + scan_chunk], spec)
            Y.append(target(s_tr.reshape(-1), spec.n_turns, gamma))
            t_gen["total"] += t["total"]
            t_gen["track"] += t["track"]
            del s_tr
        Y = np.concatenate(Y)
        X = features(k_train, spec)
        n_ic_train = spec.n_ic
    else:
        X, Y, t_gen = training_data(ctx, k_train, spec, train_ics, seed=seed + 1, chunk=scan_chunk, gamma=gamma)
        n_ic_train = train_ics
    say(f"training data done in {t_gen['total']:.1f} s; training {epochs} epochs on {len(Y)} samples")
    t0 = time.perf_counter()
    net = Net(ctx, [N_KNOBS + 3, 256, 256, 256, 256, 1], seed=seed % 1000)
    losses = train(net, X, Y, epochs=epochs, batch=batch, lr=lr)
    del X, Y
    t_train = time.perf_counter() - t0
    say(f"trained in {t_train:.1f} s (loss {losses[0]:.4g} -> {losses[-1]:.4g}); plain scan of {n_test} lattices")
    s_te, a_plain, t_plain = plain_scan(ctx, k_test, spec)
    plain_runs = np.cumprod(s_te >= spec.n_turns, axis=-1).sum(-1)
    del s_te
    if modes is None:
        modes = [dict(window=window, probe_up=probe_up, probe_down=probe_down, max_rounds=max_rounds)]
    plain_s = t_plain["total"]
    out_modes = []
    say(f"plain scan done in {t_plain['total']:.1f} s")
    for m in modes:
        runs, a_sur, t_sur, n_tracked = surrogate_scan(ctx, net, k_test, spec, band=band, workers=workers,
                                                       refine=refine, **m)
        say(f"surrogate scan {m} done in {t_sur['total']:.2f} s")
        sur_s = t_sur["total"]
        per_lat_gain = plain_s / n_test - sur_s / n_test
        rec = {
            "surrogate_s": sur_s, "surrogate_lattices_per_hour": n_test / sur_s * 3600, "speedup": plain_s / sur_s,
            "ics_tracked_surrogate": int(n_tracked), "rays_open": t_sur["n_open"] / runs.size,
            "pred_boundary_err_mean": float(np.mean(np.abs(t_sur["kb"] - plain_runs))),
            "pred_boundary_exact": float(np.mean(t_sur["kb"] == plain_runs)),
            "pred_boundary_err_q90": float(np.quantile(np.abs(t_sur["kb"] - plain_runs), 0.9)),
            "pred_boundary_within_band": float(np.mean(np.abs(t_sur["kb"] - plain_runs)
                                                       <= np.maximum(1, band * plain_runs))),
            "rounds": [[int(r[0]), round(r[1], 4), float(r[2]), round(r[3], 4)] for r in t_sur["rounds"]],
            "predict_s": t_sur["predict"], "pet_surrogate": float(sum(r[2] for r in t_sur["rounds"])),
            "timeline": t_sur["timeline"],
            "break_even_lattices": (t_gen["total"] + t_train) / per_lat_gain if per_lat_gain > 0 else None,
        }
        rec.update(area_errors(runs, plain_runs, spec))
        out_modes.append(rec)
    inf = inference_throughput(net)
    res = {
        "n_train": n_train, "n_test": n_test, "ics_per_lattice": spec.n_ic, "train_ics_per_lattice": n_ic_train,
        "train_samples": n_train * n_ic_train, "train_scan_s": t_gen["track"], "train_data_s": t_gen["total"],
        "train_s": t_train, "loss_first": losses[0], "loss_last": losses[-1],
        "plain_s": plain_s, "plain_lattices_per_hour": n_test / plain_s * 3600,
        "ics_tracked_plain": n_test * spec.n_ic, "pet_plain": t_plain["pet"],
        "mlp_forward_tflops": inf["forward_tflops"], "mlp_train_step_tflops": inf["train_tflops"],
        "mlp_forward_fused_tflops": inf.get("forward_fused_tflops"),
    }
    res.update(out_modes[0])
    if len(out_modes) > 1:
        res["modes"] = out_modes
    return res
