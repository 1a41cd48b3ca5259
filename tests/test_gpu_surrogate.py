"""GPU tests of the surrogate kernels (bmx_mlp.hip through the libbmx C ABI)
against the CPU oracle (oracle/mlp_oracle.py) and FP32/FP64 references.

Tolerances: a GEMM of BF16 operands with FP32 accumulation agrees with the
FP64 product of the same BF16 values to 1e-5 of the largest |C| (FP32
accumulation order); fused epilogues to 1e-5 relative plus one BF16 ulp on the
BF16 outputs.  Whole-network gradients agree with the oracle (which rounds to
BF16 at the same points) to 1e-3 in relative Frobenius norm per tensor: the
residue is FP32-vs-FP64 accumulation noise plus the occasional BF16 rounding
that falls the other way; a transposed operand or dropped term is O(1).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from bmx import surrogate
import mlp_oracle as mo

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _bf(a):
    return torch.from_numpy(np.asarray(a, dtype=np.float32)).to(torch.bfloat16).to(DEV)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (100, 70, 50), (1, 256, 333), (257, 3, 1), (130, 256, 35),
                                   (256, 256, 8192), (64, 128, 4096), (1, 256, 5000)])
def test_gemm_bf16(ctx, ta, tb, M, N, K):
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    a = mo.bf16(rng.standard_normal((K, M) if ta else (M, K)))
    b = mo.bf16(rng.standard_normal((N, K) if tb else (K, N)))
    A, B = _bf(a), _bf(b)
    Cf = torch.full((M, N), np.nan, dtype=torch.float32, device=DEV)
    Cb = torch.zeros((M, N), dtype=torch.bfloat16, device=DEV)
    torch.cuda.synchronize()
    lda = a.shape[1]
    ldb = b.shape[1]
    rc = ctx.lib.bmx_gemm_bf16(ctx.h, ta, tb, M, N, K, _p(A), lda, _p(B), ldb, _p(Cf), _p(Cb), N, None, 0, None)
    assert rc == 0, ctx.lib.bmx_last_error(ctx.h)
    assert ctx.lib.bmx_ctx_sync(ctx.h) == 0
    want = (a.T if ta else a) @ (b.T if tb else b)
    got = Cf.cpu().numpy().astype(np.float64)
    scale = np.max(np.abs(want)) + 1e-30
    err = np.max(np.abs(got - want)) / scale
    print(f"gemm ta={ta} tb={tb} {M}x{N}x{K}: {err:.2e}", flush=True)
    assert err < 1e-5
    assert np.array_equal(Cb.float().cpu().numpy().astype(np.float64), mo.bf16(got))


@pytest.mark.parametrize("epi", [1, 2, 3])
def test_gemm_epilogues(ctx, epi):
    M, N, K = 150, 96, 64
    rng = np.random.default_rng(epi)
    a = mo.bf16(rng.standard_normal((M, K)) * 0.3)
    b = mo.bf16(rng.standard_normal((N, K)) * 0.3)
    bias = rng.standard_normal(N).astype(np.float32)
    aux = (rng.standard_normal((M, N)) * 2).astype(np.float32)
    A, B = _bf(a), _bf(b)
    tb = torch.from_numpy(bias).to(DEV)
    taux = torch.from_numpy(aux).to(DEV)
    Cf = torch.zeros((M, N), dtype=torch.float32, device=DEV)
    Cb = torch.zeros((M, N), dtype=torch.bfloat16, device=DEV)
    torch.cuda.synchronize()
    rc = ctx.lib.bmx_gemm_bf16(ctx.h, 0, 1, M, N, K, _p(A), K, _p(B), K, _p(Cf), _p(Cb), N, _p(tb), epi, _p(taux))
    assert rc == 0 and ctx.lib.bmx_ctx_sync(ctx.h) == 0
    acc = a @ b.T
    if epi == 1:
        want_f = acc + bias
        want_b = mo.bf16(want_f)
    elif epi == 2:
        want_f = acc + bias
        want_b = mo.bf16(mo.silu(want_f))
    else:
        want_f = acc * mo.silu_grad(aux.astype(np.float64))
        want_b = mo.bf16(want_f)
    got_f = Cf.cpu().numpy().astype(np.float64)
    got_b = Cb.float().cpu().numpy().astype(np.float64)
    assert np.max(np.abs(got_f - want_f)) <= 1e-5 * np.max(np.abs(want_f))
    # BF16 outputs: within one BF16 ulp (2^-8 relative) of the FP64 value
    assert np.all(np.abs(got_b - want_b) <= 2.0 ** -7 * np.abs(want_b) + 1e-30)


def test_colsum_mse_adam(ctx):
    rng = np.random.default_rng(9)
    X = rng.standard_normal((3000, 77)).astype(np.float32)
    tX = torch.from_numpy(X).to(DEV)
    out = torch.zeros(77, dtype=torch.float32, device=DEV)
    torch.cuda.synchronize()
    assert ctx.lib.bmx_colsum_f32(ctx.h, 3000, 77, _p(tX), 77, _p(out)) == 0
    assert ctx.lib.bmx_ctx_sync(ctx.h) == 0
    assert np.allclose(out.cpu().numpy(), X.astype(np.float64).sum(0), rtol=1e-5, atol=1e-4)
    # the two-pass path at the network's shapes (16384 x 256), with more than 256 columns and a row
    # stride wider than the matrix, and the narrow path (one column; five of a wider stride)
    for M, N, ld in [(16384, 256, 256), (1000, 300, 320), (5000, 1, 1), (2000, 5, 7)]:
        Xw = rng.standard_normal((M, ld)).astype(np.float32)
        tXw = torch.from_numpy(Xw).to(DEV)
        outw = torch.zeros(N, dtype=torch.float32, device=DEV)
        torch.cuda.synchronize()
        assert ctx.lib.bmx_colsum_f32(ctx.h, M, N, _p(tXw), ld, _p(outw)) == 0
        assert ctx.lib.bmx_ctx_sync(ctx.h) == 0
        want = Xw[:, :N].astype(np.float64).sum(0)
        assert np.allclose(outw.cpu().numpy(), want, rtol=1e-5, atol=2e-4 * np.sqrt(M / 3000)), (M, N, ld)

    n = 5000
    p = rng.standard_normal(n).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    tp, tt = torch.from_numpy(p).to(DEV), torch.from_numpy(t).to(DEV)
    d = torch.zeros(n, dtype=torch.float32, device=DEV)
    db = torch.zeros(n, dtype=torch.bfloat16, device=DEV)
    loss = torch.zeros(1, dtype=torch.float32, device=DEV)
    torch.cuda.synchronize()
    assert ctx.lib.bmx_mse_grad(ctx.h, n, _p(tp), _p(tt), _p(d), _p(db), _p(loss)) == 0
    assert ctx.lib.bmx_ctx_sync(ctx.h) == 0
    e = p.astype(np.float64) - t
    assert abs(loss.item() - np.mean(e * e)) < 1e-5 * np.mean(e * e)
    assert np.allclose(d.cpu().numpy(), 2 * e / n, rtol=1e-6, atol=1e-12)

    w = rng.standard_normal(1000).astype(np.float32)
    wd = torch.from_numpy(w).to(DEV)
    wb = torch.zeros(1000, dtype=torch.bfloat16, device=DEV)
    m = torch.zeros(1000, dtype=torch.float32, device=DEV)
    v = torch.zeros(1000, dtype=torch.float32, device=DEV)
    w64, m64, v64 = w.astype(np.float64), np.zeros(1000), np.zeros(1000)
    for step in range(1, 4):
        g = rng.standard_normal(1000).astype(np.float32)
        gd = torch.from_numpy(g).to(DEV)
        torch.cuda.synchronize()
        assert ctx.lib.bmx_adam_step(ctx.h, 1000, _p(wd), _p(wb), _p(gd), _p(m), _p(v), 1e-2, 0.9, 0.999, 1e-8,
                                     step) == 0
        assert ctx.lib.bmx_ctx_sync(ctx.h) == 0
        mo.adam(w64, g.astype(np.float64), m64, v64, 1e-2, 0.9, 0.999, 1e-8, step)
    got = wd.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(got - w64)) < 1e-5
    assert np.array_equal(wb.float().cpu().numpy().astype(np.float64), mo.bf16(got))


@pytest.mark.parametrize("n", [1000, 4096])
def test_network_forward_backward_vs_oracle(ctx, n):
    dims = [35, 256, 256, 256, 256, 1]
    params = mo.init_params(dims, seed=5)
    net = surrogate.MLP(ctx, dims, params=params)
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, (n, 35)).astype(np.float32)
    t = rng.uniform(0, 1, n).astype(np.float32)
    tx, tt = torch.from_numpy(x).to(DEV), torch.from_numpy(t).to(DEV)
    torch.cuda.synchronize()
    y = net.forward(tx)
    loss = net.backward(tt)
    net.sync()
    p32 = [(W.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)) for W, b in params]
    want_loss, want_g, want_y = mo.loss_and_grads(p32, x.astype(np.float64), t.astype(np.float64))
    got_y = y.cpu().numpy().astype(np.float64)
    ey = np.max(np.abs(got_y - want_y)) / np.max(np.abs(want_y))
    el = abs(loss.item() - want_loss) / want_loss
    print(f"forward {ey:.2e} loss {el:.2e}", flush=True)
    # the device's SiLU takes the hardware exponential and reciprocal (about
    # 1 ulp each): where that moves a pre-activation's SiLU across a BF16
    # rounding boundary, the oracle's BF16 emulation rounds the other way,
    # one BF16 ulp (2^-8) on that activation, which reaches the output at
    # about 1e-3 of its range
    assert ey < 3e-3 and el < 1e-4
    for li, ((gW, gb), (wW, wb)) in enumerate(zip(net.grads(), want_g)):
        eW = np.linalg.norm(gW - wW) / np.linalg.norm(wW)
        eb = np.linalg.norm(gb - wb) / np.linalg.norm(wb)
        print(f"layer {li}: dW {eW:.2e} db {eb:.2e}", flush=True)
        assert eW < 1e-3 and eb < 1e-3, (li, eW, eb)


def test_training_reduces_loss(ctx):
    """A few hundred Adam steps on a smooth target halve the loss at least."""
    dims = [3, 64, 64, 1]
    net = surrogate.MLP(ctx, dims, seed=1)
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (2048, 3)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) * np.cos(x[:, 1]) + 0.5 * x[:, 2] ** 2).astype(np.float32)
    tx, tt = torch.from_numpy(x).to(DEV), torch.from_numpy(t).to(DEV)
    torch.cuda.synchronize()
    first = None
    for i in range(300):
        loss = net.train_step(tx, tt, 3e-3)
        if i == 0:
            net.sync()
            first = loss.item()
    net.sync()
    last = loss.item()
    print(f"loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.5 * first


def test_da_features_match_host_rows(ctx):
    """bmx_da_features forms, in BF16, exactly the rows surrogate.features
    builds on the host."""
    spec = surrogate.ScanSpec(n_turns=64, n_angle=5, n_amp=7, n_pz=3)
    rng = np.random.default_rng(3)
    knobs = rng.uniform(-0.2, 0.2, (4, surrogate.N_KNOBS))
    want = mo.bf16(surrogate.features(knobs, spec))
    geo = torch.from_numpy(surrogate.geometry(spec).astype(np.float32)).to(DEV)
    kd = torch.from_numpy(knobs.astype(np.float32)).to(DEV)
    X = torch.zeros(4 * spec.n_ic, surrogate.N_KNOBS + 3, dtype=torch.bfloat16, device=DEV)
    torch.cuda.synchronize()
    rc = ctx.lib.bmx_da_features(ctx.h, 4, spec.n_ic, surrogate.N_KNOBS, 3, _p(kd), surrogate.KNOB_RANGE, _p(geo),
                                 _p(X))
    assert rc == 0 and ctx.lib.bmx_ctx_sync(ctx.h) == 0
    got = X.float().cpu().numpy().astype(np.float64)
    # the host rows are float32 knobs / 0.2; the device scales float32 knobs by 1 / 0.2
    assert np.max(np.abs(got - want)) <= 2.0 ** -8 * np.max(np.abs(want))


@pytest.mark.parametrize("K", [1, 7, 128, 300])
def test_step_fit_matches_host(ctx, K):
    """bmx_step_fit picks, for every ray, a boundary whose least-squares step
    cost equals the host step_fit_boundary's minimum (the sums are FP64 in
    both, accumulated in different orders, so a near-tie may resolve to a
    different k of equal cost), and the same k where the minimum is clear."""
    rng = np.random.default_rng(K)
    n = 3000
    # noisy steps at random positions, some flat profiles (ties), some all-survive / all-lost
    kt = rng.integers(0, K + 1, n)
    p = np.where(np.arange(K)[None, :] < kt[:, None], 1.0, 0.2) + 0.05 * rng.standard_normal((n, K))
    p[:50] = 1.0
    p[50:100] = 0.0
    p = p.astype(np.float32)
    want = surrogate.step_fit_boundary(p, 1.0)
    tp = torch.from_numpy(p).to(DEV)
    out = torch.full((n,), -1, dtype=torch.int32, device=DEV)
    torch.cuda.synchronize()
    assert ctx.lib.bmx_step_fit(ctx.h, n, K, _p(tp), 1.0, _p(out)) == 0
    assert ctx.lib.bmx_ctx_sync(ctx.h) == 0
    got = out.cpu().numpy().astype(np.int64)
    assert np.all((got >= 0) & (got <= K))

    def cost(k):
        q = p.astype(np.float64)
        c = np.empty(n)
        for i in range(n):
            a, b = q[i, :k[i]], q[i, k[i]:]
            c[i] = np.sum((a - 1.0) ** 2) + (np.sum((b - b.mean()) ** 2) if len(b) else 0.0)
        return c

    cg, cw = cost(got), cost(want)
    assert np.all(cg <= cw + 1e-9 * (1 + np.abs(cw)))
    same = got == want
    print(f"K={K}: identical k on {same.mean():.4f} of rays", flush=True)
    assert same.mean() > 0.99
    assert np.all(got[:50] == K) and np.all(want[:50] == K)


@pytest.mark.parametrize("dims,n", [([35, 256, 256, 256, 256, 1], 5000), ([35, 256, 256, 256, 256, 1], 64),
                                    ([3, 256, 1], 1000), ([64, 256, 256, 1], 777)])
def test_fused_forward_equals_layered(ctx, dims, n):
    """bmx_mlp_forward (all layers in one kernel, activations in LDS) gives
    the layer-by-layer pass's outputs bit for bit: same MFMA operands in the
    same k order, same epilogues and BF16 roundings."""
    net = surrogate.MLP(ctx, dims, seed=3)
    rng = np.random.default_rng(4)
    xb = torch.from_numpy(rng.uniform(-1, 1, (n, dims[0])).astype(np.float32)).to(torch.bfloat16).to(DEV)
    torch.cuda.synchronize()
    y = net.forward(xb)
    net.sync()  # the library's stream has written y before torch copies it
    want = y.clone()
    torch.cuda.synchronize()
    got = net.forward_fused(xb)
    net.sync()
    a, b = got.cpu().numpy(), want.cpu().numpy()
    print(f"fused vs layered: max |dy| {np.max(np.abs(a - b)):.3e}", flush=True)
    assert np.array_equal(a, b)


def test_fused_forward_rejects_other_shapes(ctx):
    net = surrogate.MLP(ctx, [3, 64, 64, 1], seed=1)
    assert not net.fused_ok()
    xb = torch.zeros(8, 3, dtype=torch.bfloat16, device=DEV)
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="bmx_mlp_forward"):
        net.forward_fused(xb)


def test_da_screen_equals_layered_screening(ctx):
    """bmx_da_screen (inputs formed in the kernel, fused forward, step fit)
    gives the boundaries of bmx_da_features + layer-by-layer forward +
    bmx_step_fit, over several lattice chunks."""
    spec = surrogate.ScanSpec(n_turns=64, n_angle=16, n_amp=128, n_pz=5)
    net = surrogate.MLP(ctx, [35, 256, 256, 256, 256, 1], seed=7)
    rng = np.random.default_rng(9)
    knobs = rng.uniform(-0.2, 0.2, (300, surrogate.N_KNOBS))
    a = surrogate.predict_boundaries(net, knobs, spec)
    b = surrogate.predict_boundaries(net, knobs, spec, lat_chunk=97, layered=True)
    assert a.shape == (300, 5, 16)
    assert np.array_equal(a, b)


def _toy_data(n, d0, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, d0)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) * np.cos(x[:, 1 % d0]) + 0.5 * x[:, 2 % d0] ** 2).astype(np.float32)
    return x, t


@pytest.mark.parametrize("dims", [[35, 256, 256, 256, 256, 1], [3, 64, 64, 1]])
def test_library_network_graph_training_equals_eager(ctx, dims):
    """bmx_surrogate_train replaying each step as a hipGraph gives the
    eager launches' parameters bit for bit, over several epochs (host reads
    of the loss between replays), and trains: the loss falls."""
    x, t = _toy_data(8192, dims[0], 11)
    a = surrogate.Net(ctx, dims, seed=2)
    b = surrogate.Net(ctx, dims, seed=2)
    la = a.train(x, t, epochs=3, batch=512, lr=3e-3, lr_final=3e-4, seed=5)
    lb = b.train(x, t, epochs=3, batch=512, lr=3e-3, lr_final=3e-4, seed=5, eager=True)
    print("graph losses", la, "eager losses", lb, flush=True)
    assert la == lb
    for (Wa, ba), (Wb, bb) in zip(a.params(), b.params()):
        assert np.array_equal(Wa, Wb) and np.array_equal(ba, bb)
    assert la[-1] < 0.7 * la[0]
    a.close()
    b.close()


def test_library_network_trains_like_the_python_loop(ctx):
    """Same network, data and schedule through the C++ loop and the Python
    loop (different shuffles): held-out errors within 25 % of each other."""
    dims = [3, 64, 64, 1]
    x, t = _toy_data(16384, 3, 12)
    xv, tv = _toy_data(4096, 3, 13)
    cnet = surrogate.Net(ctx, dims, seed=4)
    pnet = surrogate.MLP(ctx, dims, seed=4)
    surrogate.train(cnet, x, t, epochs=6, batch=256, lr=3e-3, lr_final=3e-4, seed=1)
    surrogate.train(pnet, x, t, epochs=6, batch=256, lr=3e-3, lr_final=3e-4, seed=1)
    errs = []
    for p in (cnet.params(), pnet.params()):
        h = xv.astype(np.float64)
        for W, b in p[:-1]:
            z = h @ W.T + b
            h = z / (1 + np.exp(-z))
        y = (h @ p[-1][0].T + p[-1][1])[:, 0]
        errs.append(float(np.mean((y - tv) ** 2)))
    print("held-out mse C++ loop %.3e, Python loop %.3e" % tuple(errs), flush=True)
    assert errs[0] < 1.25 * errs[1] and errs[1] < 1.25 * errs[0]
    assert errs[0] < 0.05 * np.var(tv)
    cnet.close()


@pytest.mark.parametrize("dims", [[35, 256, 256, 256, 256, 1], [3, 64, 64, 1]])
def test_library_network_infer_and_screen(ctx, dims):
    """bmx_surrogate_infer and bmx_surrogate_screen equal the Python
    network's passes with the same parameters bit for bit."""
    pnet = surrogate.MLP(ctx, dims, seed=6)
    cnet = surrogate.Net(ctx, dims, params=pnet.params())
    rng = np.random.default_rng(14)
    xb = torch.from_numpy(rng.uniform(-1, 1, (3000, dims[0])).astype(np.float32)).to(torch.bfloat16).to(DEV)
    torch.cuda.synchronize()
    want = pnet.forward(xb)
    pnet.sync()
    want = want.cpu().numpy()
    got = cnet.infer(xb).cpu().numpy()
    assert np.array_equal(got, want)
    if dims[0] == 35:
        spec = surrogate.ScanSpec(n_turns=64, n_angle=8, n_amp=64, n_pz=3)
        knobs = rng.uniform(-0.2, 0.2, (40, surrogate.N_KNOBS))
        assert np.array_equal(surrogate.predict_boundaries(cnet, knobs, spec),
                              surrogate.predict_boundaries(pnet, knobs, spec, layered=True))
    cnet.close()
