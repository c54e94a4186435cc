"""GPU kernels against the FP64 oracle on seeded random problems.

Tolerances (normwise rel-L2 = ||gpu - oracle|| / ||oracle||):
  FP64 1e-12;  FP32 1e-5;  BF16 / FP16 inputs with FP32 output, against the
  oracle evaluated on the same rounded inputs: 1e-5 (accumulation only);
  BF16 / FP16 end to end against the unrounded FP64 oracle: 1e-2.
Pooling is compared bit for bit with the oracle's restatement of Owl's C."""
import numpy as np
import pytest

import oracle_ndarray as on
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID, conv_out_dim, k_eff

torch = pytest.importorskip("torch")

DEV = "cuda:0"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def problem(seed, nd_, aligned):
    rng = np.random.default_rng(seed)
    sp = [int(rng.integers(5, 14)) for _ in range(nd_)]
    ksp = [int(rng.choice([1, 2, 3])) for _ in range(nd_)]
    stride = [int(rng.integers(1, 3)) for _ in range(nd_)]
    rate = [int(rng.integers(1, 3)) for _ in range(nd_)]
    pad = [SAME, VALID][seed % 2]
    if aligned:
        Ci, Co = int(rng.choice([8, 16, 24, 64])), int(rng.choice([8, 16, 40, 128]))
    else:
        Ci, Co = int(rng.integers(1, 13)), int(rng.integers(1, 13))
    B = int(rng.integers(1, 4))
    for i in range(nd_):
        while conv_out_dim(pad, sp[i], k_eff(ksp[i], rate[i]), stride[i]) <= 0:
            sp[i] += 2
    x = rng.standard_normal([B, *sp, Ci])
    w = rng.standard_normal([*ksp, Ci, Co]) / np.sqrt(np.prod(ksp) * Ci)
    return x, w, stride, rate, pad


def t(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(dtype)


def run3(x, w, stride, rate, pad, dtype, out_dtype=None):
    f = nd.dilated_conv3d if x.ndim == 5 else nd.dilated_conv2d
    fb_i = nd.dilated_conv3d_backward_input if x.ndim == 5 else nd.dilated_conv2d_backward_input
    fb_k = nd.dilated_conv3d_backward_kernel if x.ndim == 5 else nd.dilated_conv2d_backward_kernel
    xt, wt = t(x, dtype), t(w, dtype)
    y = f(xt, wt, stride, rate, pad, out_dtype=out_dtype)
    rng = np.random.default_rng(7)
    g = rng.standard_normal(tuple(y.shape))
    gt = t(g, dtype)
    dx = fb_i(xt, wt, stride, rate, gt, out_dtype=out_dtype)
    dw = fb_k(xt, wt, stride, rate, gt, out_dtype=out_dtype)
    torch.cuda.synchronize()
    c = lambda a: a.float().cpu().numpy().astype(np.float64) if a.dtype != torch.float64 else a.cpu().numpy()
    return c(y), c(dx), c(dw), g


def oracle3(x, w, stride, rate, pad, g):
    y = on.dilated_conv2d(x, w, stride, rate, pad)
    return y, on.dilated_conv2d_backward_input(x, w, stride, rate, g), \
        on.dilated_conv2d_backward_kernel(x, w, stride, rate, g)


CASES = [(s, d, a) for s in range(10) for d in (2, 3) for a in (True, False)]


@pytest.mark.gpu
@pytest.mark.parametrize("seed,nd_,aligned", CASES)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_conv_fp32_fp64_vs_oracle(seed, nd_, aligned, dtype):
    x, w, stride, rate, pad = problem(seed, nd_, aligned)
    td = torch.float64 if dtype == "f64" else torch.float32
    if dtype == "f32":
        x, w = x.astype(np.float32).astype(np.float64), w.astype(np.float32).astype(np.float64)
    y, dx, dw, g = run3(x, w, stride, rate, pad, td)
    if dtype == "f32":
        g = g.astype(np.float32).astype(np.float64)
    ry, rdx, rdw = oracle3(x, w, stride, rate, pad, g)
    tol = 1e-12 if dtype == "f64" else 1e-5
    assert rel(y, ry) < tol and rel(dx, rdx) < tol and rel(dw, rdw) < tol, (rel(y, ry), rel(dx, rdx), rel(dw, rdw))


@pytest.mark.gpu
@pytest.mark.parametrize("seed,nd_,aligned", CASES[::2])
@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_conv_half_precision_vs_oracle(seed, nd_, aligned, dtype):
    x, w, stride, rate, pad = problem(seed, nd_, aligned)
    td = torch.bfloat16 if dtype == "bf16" else torch.float16
    rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
    # FP32 output: the MFMA arithmetic itself against the oracle on rounded inputs
    y, dx, dw, g = run3(x, w, stride, rate, pad, td, out_dtype=torch.float32)
    ry, rdx, rdw = oracle3(rnd(x), rnd(w), stride, rate, pad, rnd(g))
    assert rel(y, ry) < 1e-5 and rel(dx, rdx) < 1e-5 and rel(dw, rdw) < 1e-5
    # storage-dtype output end to end against the unrounded FP64 oracle
    y, dx, dw, g = run3(x, w, stride, rate, pad, td)
    ry, rdx, rdw = oracle3(x, w, stride, rate, pad, g)
    assert rel(y, ry) < 1e-2 and rel(dx, rdx) < 1e-2 and rel(dw, rdw) < 1e-2


def _same_bits(a, b):
    return np.array_equal(a, b, equal_nan=True) and np.array_equal(np.signbit(a), np.signbit(b))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pool_bit_exact(seed, dt):
    rng = np.random.default_rng(seed)
    nd_ = 2 + seed % 2
    sp = [int(rng.integers(4, 12)) for _ in range(nd_)]
    C = int(rng.choice([1, 3, 4, 16]))
    k = [int(rng.integers(1, 4)) for _ in range(nd_)]
    s = [int(rng.integers(1, 4)) for _ in range(nd_)]
    x = rng.standard_normal([2, *sp, C]).astype(dt)
    x[x > 1.5] = x.max()  # ties
    for pad in (SAME, VALID):
        if any(conv_out_dim(pad, n, kk, ss) <= 0 for n, kk, ss in zip(sp, k, s)):
            continue
        names = ["max_pool%dd" % nd_, "avg_pool%dd" % nd_]
        for name in names:
            yg = getattr(nd, name)(t(x, torch.float32 if dt == np.float32 else torch.float64), k, s, pad)
            yo = getattr(on, name)(x, k, s, pad)
            assert _same_bits(yg.cpu().numpy(), yo), name
            dy = rng.standard_normal(yo.shape).astype(dt)
            dxg = getattr(nd, name + "_backward")(pad, t(x, yg.dtype), k, s, t(dy, yg.dtype))
            dxo = getattr(on, name + "_backward")(pad, x, k, s, dy)
            assert _same_bits(dxg.cpu().numpy(), dxo), name + "_backward"
            # host drop-in path gives the same bits
            assert _same_bits(getattr(nd, name)(x, k, s, pad), yo)
            assert _same_bits(getattr(nd, name + "_backward")(pad, x, k, s, dy), dxo)
        if nd_ == 2:
            yg, ag = nd.max_pool2d_argmax(t(x, torch.float32 if dt == np.float32 else torch.float64), k, s, pad)
            yo, ao = on.max_pool2d_argmax(x, k, s, pad)
            assert _same_bits(yg.cpu().numpy(), yo)
            assert np.array_equal(ag.cpu().numpy(), ao)
            yh, ah = nd.max_pool2d_argmax(x, k, s, pad)
            assert _same_bits(yh, yo) and np.array_equal(ah, ao)


@pytest.mark.gpu
def test_pool_nan_windows_and_flat_index_zero():
    x = np.array([1, 3, 3, 2,
                  np.nan, np.nan, 5, 0,
                  np.nan, np.nan, 0, 0,
                  np.nan, np.nan, 0, 7], np.float32).reshape(1, 4, 4, 1)
    dy = np.array([1, 2, 4, 8], np.float32).reshape(1, 2, 2, 1)
    y, a = nd.max_pool2d_argmax(x, [2, 2], [2, 2], VALID)
    assert a.ravel().tolist() == [1, 6, -1, 15]
    dx = nd.max_pool2d_backward(VALID, x, [2, 2], [2, 2], dy)
    assert _same_bits(dx, on.max_pool2d_backward(VALID, x, [2, 2], [2, 2], dy))
    assert dx.ravel()[0] == 4
    # the flag is cleared: a clean call right after gives zero at index 0
    x2 = np.arange(16, dtype=np.float32).reshape(1, 4, 4, 1)
    dx2 = nd.max_pool2d_backward(VALID, x2, [2, 2], [2, 2], dy)
    assert _same_bits(dx2, on.max_pool2d_backward(VALID, x2, [2, 2], [2, 2], dy))


@pytest.mark.gpu
@pytest.mark.parametrize("nd_", [2, 3])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pool_backward_tiled_windows_bit_exact(nd_, dt):
    """Windows that tile the input exactly (kernel == stride, VALID, sizes a
    multiple of the stride) take the one-thread-per-window backward kernels;
    the results keep the oracle's bits, including -0 and NaN gradients (Owl
    adds them to 0: -0 becomes +0), with 4-wide and odd channel counts."""
    rng = np.random.default_rng(11 + nd_)
    for k, C in (([2] * nd_, 16), ([3, 2, 2][:nd_], 5), ([1, 3, 2][:nd_], 8)):
        sp = [kk * int(rng.integers(2, 5)) for kk in k]
        x = rng.standard_normal([2, *sp, C]).astype(dt)
        x[0, (0,) * nd_] = np.nan
        for name in ("max_pool%dd" % nd_, "avg_pool%dd" % nd_):
            yo = getattr(on, name)(x, k, k, VALID)
            dy = rng.standard_normal(yo.shape).astype(dt)
            dy.ravel()[::7] = -0.0
            dy.ravel()[3] = np.nan
            td = torch.float32 if dt == np.float32 else torch.float64
            dxg = getattr(nd, name + "_backward")(VALID, t(x, td), k, k, t(dy, td))
            dxo = getattr(on, name + "_backward")(VALID, x, k, k, dy)
            assert _same_bits(dxg.cpu().numpy(), dxo), (name, k, C)
