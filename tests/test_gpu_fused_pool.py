"""The conv layer followed by a 2x2 stride-2 VALID max-pool layer in one call
(owl_amd_conv_fwd_bias_act_pool), the end of a VGG block.

For BF16 / FP16 stride-1 layers the halo kernel pools its staged output tile
before storing it; otherwise the convolution runs, then the pooling kernel.
Either way:
  * pooled equals Owl's max_pool2d VALID [2;2] [2;2] of y bit for bit: the
    FP64 oracle's pooling of y (owl_ndarray_pool_impl.h order, first strict
    maximum, NaN never taken, -inf for an all-NaN window) and the device
    pooling kernel over y;
  * y and pre equal the unfused conv + bias + activation bit for bit when
    both run the same MFMA fragment edge (tile rows do not change any
    output's summation), and y is within 1e-2 (16-bit output) of the FP64
    oracle on the rounded inputs;
  * without y (inference) the pooled result is the same;
  * y is the same whether or not pre is kept (keeping it, the 16-bit
    epilogue forms y from the staged pre-activation).
Shapes cover 224- and 448-pixel tiles, odd output heights and widths (the
last row / column is dropped), a partial channel slice, a layer whose tiles
are one row (the pooling kernel runs apart) and NaN windows."""
import numpy as np
import pytest

import oracle_ndarray as on
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")

DEV = "cuda:0"
TD = {"bf16": torch.bfloat16, "f16": torch.float16}

# (x shape, kernel shape, padding, dtype, activation, NaNs in x)
CASES = [
    ([2, 56, 56, 64], [3, 3, 64, 128], SAME, "bf16", "relu", False),   # VGG layer: 224-pixel tiles, 4 rows
    ([2, 56, 56, 128], [3, 3, 128, 64], SAME, "bf16", "relu", False),  # 64-wide slices: 448-pixel tiles
    ([3, 15, 13, 32], [3, 3, 32, 72], SAME, "bf16", "relu", False),    # odd Ho, Wo; partial channel slice
    ([2, 29, 30, 64], [3, 3, 64, 128], VALID, "f16", "relu", False),   # Ho 27: last odd row dropped
    ([1, 20, 120, 32], [3, 3, 32, 128], SAME, "bf16", "none", False),  # Wo 120: 2-row tiles (32-edge only)
    ([1, 9, 200, 32], [3, 3, 32, 128], SAME, "bf16", "relu", False),   # one-row tiles: pooling kernel apart
    ([2, 28, 28, 64], [1, 1, 64, 64], SAME, "f16", "relu", False),     # 1x1 kernel
    ([2, 14, 18, 32], [3, 3, 32, 64], SAME, "bf16", "none", True),     # NaN windows, no activation
    ([2, 14, 18, 32], [3, 3, 32, 64], SAME, "f16", "relu", True),      # NaN windows under relu
]


def dev(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(dt)


def host(t):
    return t.double().cpu().numpy()


def bits(t):
    return t.view(torch.int16)


def pool_tiles_fit(ms, Ho, Wo, N, kh, kw):
    """Whether the halo kernel's fused pooling takes fragment edge ms: whole
    rows per tile, an even count unless one tile covers the image, and an
    input halo within the two 352-pixel LDS buffers."""
    cap = (448 if ms == 16 else 512) if N <= 64 else (224 if ms == 16 else 256)
    tr = min(cap // Wo, Ho)
    if tr < Ho:
        tr &= ~1
    return tr >= 1 and (tr + kh - 1) * (Wo + kw - 1) <= 704


def rel(a, b):
    ok = np.isfinite(b)
    return np.linalg.norm(a[ok] - b[ok]) / max(np.linalg.norm(b[ok]), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("edge", [None, "16", "32"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_conv_bias_act_maxpool(case, edge, monkeypatch):
    xs, ks, pad, dtn, act, nans = CASES[case]
    if edge is not None:
        monkeypatch.setenv("OWL_AMD_HALO_MS", edge)
    print(f"case {case} x {xs} k {ks} {dtn} act {act} edge {edge}", flush=True)
    rng = np.random.default_rng(70 + case)
    x = rng.standard_normal(xs)
    if nans:
        x.reshape(-1, xs[-1])[rng.choice(x.size // xs[-1], x.size // xs[-1] // 6, replace=False)] = np.nan
    w = rng.standard_normal(ks) / np.sqrt(np.prod(ks[:-1]))
    bias = rng.standard_normal(ks[-1]) * 0.5
    td = TD[dtn]
    xd, wd, bd = dev(x, td), dev(w, td), torch.from_numpy(bias).float().to(DEV)

    y, pooled, pre = nd.conv2d_bias_act_maxpool(xd, wd, bd, [1, 1], pad, act=act, keep_pre=True)
    y_ref, pre_ref = nd.conv2d_bias_act(xd, wd, bd, [1, 1], pad, act=act, keep_pre=True)
    y_plain = nd.conv2d_bias_act(xd, wd, bd, [1, 1], pad, act=act)
    none_y, pooled2 = nd.conv2d_bias_act_maxpool(xd, wd, bd, [1, 1], pad, act=act, keep_y=False)
    pooled_dev = nd.max_pool2d(y, [2, 2], [2, 2], VALID)
    torch.cuda.synchronize()

    Ho, Wo = y.shape[1], y.shape[2]
    assert none_y is None
    assert list(pooled.shape) == [xs[0], Ho // 2, Wo // 2, ks[-1]]
    # pooled is Owl's max-pool of y, bit for bit
    assert torch.equal(bits(pooled), bits(pooled_dev))
    assert torch.equal(bits(pooled2), bits(pooled))
    ref_pool = on.max_pool2d(host(y), [2, 2], [2, 2], VALID)
    np.testing.assert_array_equal(host(pooled), ref_pool)
    # keeping pre, y is formed from the staged pre-activation: the same bits
    assert torch.equal(bits(y_plain), bits(y_ref))
    # y and pre as the unfused layer computes them
    # (the same edge runs both when the forced one takes the fused pooling,
    # or when no edge does and the plain convolution runs)
    fits = [pool_tiles_fit(ms, Ho, Wo, ks[-1], ks[0], ks[1]) for ms in (16, 32)]
    if edge is not None and (fits[int(edge) // 32] or not any(fits)):
        assert torch.equal(bits(y), bits(y_ref))
        assert torch.equal(bits(pre), bits(pre_ref))
    rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
    ref_pre = on.conv2d(rnd(x), rnd(w), [1, 1], pad) + bias.astype(np.float32).astype(np.float64)
    ref_y = np.fmax(ref_pre, 0.0) if act == "relu" else ref_pre
    assert rel(host(y), ref_y) < 1e-2
    assert rel(host(pre), ref_pre) < 1e-2
    if nans and act == "none":
        assert np.isneginf(host(pooled)).any() and not np.isnan(host(pooled)).any()
    if nans and act == "relu":
        assert not np.isnan(host(y)).any() and np.isnan(host(pre)).any()


@pytest.mark.gpu
def test_cgraph_vgg_block_pools_in_the_conv_epilogue():
    # conv + bias + relu twice, then the 2x2 max-pool: the optimised graph runs
    # the second layer and the pooling as one call without y; its result (and
    # every HIP-graph replay's) equals those ndarray calls in sequence
    from owl_amd import cgraph as G
    dt = torch.bfloat16
    gx = G.var_arr("x", [4, 28, 28, 32], dt)
    gw1, gb1 = G.var_arr("w1", [3, 3, 32, 64], dt), G.var_arr("b1", [64], dt)
    gw2, gb2 = G.var_arr("w2", [3, 3, 64, 64], dt), G.var_arr("b2", [64], dt)
    h1 = G.relu(G.add(G.conv2d(gx, gw1, [1, 1]), gb1))
    gy = G.max_pool2d(G.relu(G.add(G.conv2d(h1, gw2, [1, 1]), gb2)), [2, 2], [2, 2], VALID)
    g = G.optimise(G.make_graph([gx, gw1, gb1, gw2, gb2], [gy], "vgg_block"))
    assert "conv2d_bias_act_maxpool" in [n.op for n in g.order]

    def values(seed):
        gen = torch.Generator(device=DEV).manual_seed(seed)
        r = lambda *s, sc=1.0: (torch.randn(*s, device=DEV, generator=gen) * sc).to(dt)
        return [r(4, 28, 28, 32), r(3, 3, 32, 64, sc=0.08), r(64, sc=0.1), r(3, 3, 64, 64, sc=0.06), r(64, sc=0.1)]

    def eager(v):
        x, w1, b1, w2, b2 = v
        h = nd.conv2d_bias_act(x, w1, b1.float(), [1, 1], SAME, act="relu")
        y, p = nd.conv2d_bias_act_maxpool(h, w2, b2.float(), [1, 1], SAME, act="relu")
        assert torch.equal(bits(p), bits(nd.max_pool2d(y, [2, 2], [2, 2], VALID)))
        return p

    v = values(1)
    for n_, t_ in zip([gx, gw1, gb1, gw2, gb2], v):
        G.assign_arr(n_, t_)
    assert torch.equal(bits(G.eval_graph(g)[0]), bits(eager(v)))
    G.capture_graph(g)
    for rep in range(3):
        v = values(10 + rep)
        for n_, t_ in zip([gx, gw1, gb1, gw2, gb2], v):
            G.assign_arr(n_, t_)
        got = G.replay_graph(g)[0].clone()
        ref = eager(v)
        torch.cuda.synchronize()
        assert torch.equal(bits(got), bits(ref)), rep


BWD_TD = {"f32": torch.float32, "f64": torch.float64, "bf16": torch.bfloat16, "f16": torch.float16}
IBITS = {torch.float32: torch.int32, torch.float64: torch.int64, torch.bfloat16: torch.int16,
         torch.float16: torch.int16}


@pytest.mark.gpu
@pytest.mark.parametrize("dtn", ["f64", "f32", "bf16", "f16"])
@pytest.mark.parametrize("shape", [[2, 56, 56, 64], [3, 8, 10, 20], [2, 7, 9, 24]])
def test_maxpool_bias_act_backward(shape, dtn):
    # against max_pool2d_backward into dy then bias_act_backward: dz bit for
    # bit (odd H / W run those two calls), db within rounding and the same on
    # every call; FP64 also against the oracle's max-pool backward times Owl's
    # relu factor (exact) and the FP64 sum of dz
    print(f"maxpool_bias_act_backward {shape} {dtn}", flush=True)
    td = BWD_TD[dtn]
    B, H, W, C = shape
    rng = np.random.default_rng(90 + H + C)
    pre_h = rng.standard_normal(shape)
    flat = pre_h.reshape(-1)
    flat[rng.choice(flat.size, flat.size // 50, replace=False)] = 0.0      # factor 1/2
    npix = flat.size // C
    pre_h.reshape(npix, C)[rng.choice(npix, max(npix // 100, 1), replace=False), 0] = np.nan  # db[0] is NaN
    pre = dev(pre_h, td)
    y = torch.fmax(pre, torch.zeros((), dtype=td, device=DEV))
    dp_h = rng.standard_normal([B, H // 2, W // 2, C])
    dp_h.reshape(-1)[::97] = -0.0  # Owl adds the gradient to a zeroed dx: -0 arrives as +0
    dp = dev(dp_h, td)

    dz, db = nd.maxpool_bias_act_backward(y, pre, dp)
    dz2, db2 = nd.maxpool_bias_act_backward(y, pre, dp)
    dy = nd.max_pool2d_backward(VALID, y, [2, 2], [2, 2], dp)
    dz_ref, db_ref = nd.bias_act_backward(pre, dy, act="relu")
    torch.cuda.synchronize()
    ib, ibd = IBITS[td], IBITS[db.dtype]
    assert torch.equal(dz.view(ib), dz_ref.view(ib))
    # IEEE's -0 for a negative gradient times the factor 0, in every dtype
    negzero = (dz == 0) & torch.signbit(dz)
    assert bool(negzero.any()), "no -0 products in the test data"
    assert torch.equal(dz2.view(ib), dz.view(ib)) and torch.equal(db2.view(ibd), db.view(ibd))
    dbh, dbr = host(db), host(db_ref)
    ok = ~np.isnan(dbr)
    assert np.array_equal(np.isnan(dbh), ~ok)
    tol = 1e-12 if dtn == "f64" else 1e-5
    assert np.abs(dbh[ok] - dbr[ok]).max() <= tol * max(np.abs(dbr[ok]).max(), 1.0)
    if dtn == "f64":
        fac = (np.sign(pre_h) + 1.0) / 2.0
        ref = on.max_pool2d_backward(VALID, host(y), [2, 2], [2, 2], host(dp)) * fac
        np.testing.assert_array_equal(host(dz), ref)
        np.testing.assert_allclose(dbh[ok], ref.reshape(-1, C).sum(axis=0)[ok], rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("hw", [(8, 10), (9, 7)])
def test_algodiff_conv_bias_relu_maxpool_vs_torch_autograd(hw):
    # gradients of sum_sqr(maxpool(relu(conv(x, w) + b))) against torch's FP64
    # autograd of the same layers (NCHW); even sizes take the one-pass
    # max-pool / relu adjoint, odd ones the two calls
    from owl_amd import algodiff as ad
    H, W = hw
    rng = np.random.default_rng(5 + H)
    x = rng.standard_normal([2, H, W, 6])
    w = rng.standard_normal([3, 3, 6, 5]) / 5
    b = rng.standard_normal(5) * 0.3
    xt, wt, bt = (dev(a, torch.float64) for a in (x, w, b))
    gx, gw, gb = ad.grad(lambda a, k, bb: ad.sum_sqr(ad.conv2d_bias_relu_maxpool(a, k, bb, [1, 1], SAME)),
                         xt, wt, bt)
    X = torch.tensor(x.transpose(0, 3, 1, 2).copy(), requires_grad=True)
    Wt = torch.tensor(w.transpose(3, 2, 0, 1).copy(), requires_grad=True)
    Bt = torch.tensor(b.copy(), requires_grad=True)
    y = torch.nn.functional.max_pool2d(torch.relu(torch.nn.functional.conv2d(X, Wt, Bt, padding=1)), 2, 2)
    (y * y).sum().backward()
    r = lambda a, b_: np.linalg.norm(a - b_) / np.linalg.norm(b_)
    assert r(host(gx), X.grad.numpy().transpose(0, 2, 3, 1)) < 1e-12
    assert r(host(gw), Wt.grad.numpy().transpose(2, 3, 1, 0)) < 1e-12
    assert r(host(gb), Bt.grad.numpy()) < 1e-12
