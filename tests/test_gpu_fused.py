"""The conv layer's fused epilogue, y = relu(conv(x, w) + b), and its adjoint.

Forward: against the FP64 oracle's convolution plus the bias and numpy's
fmax (Owl's relu is fmax(v, 0), so NaN maps to 0), on shapes that take each
kernel route (16-bit stride-1 halo kernel, aligned direct-A kernel, generic
implicit GEMM for strided / unaligned / FP64 / 3-D).  Tolerances as for the
plain convolution: FP64 1e-12, FP32 1e-5, 16-bit inputs with FP32 output 1e-5
(against the oracle on the rounded inputs), 16-bit output 1e-2.  In FP32 the
fused result must equal fmax(conv + b, 0) computed from the unfused kernel's
output bit for bit (the bias is added to the same FP32 accumulator).

Backward: dz = dy * (signum(pre) + 1) / 2 holds exactly (the factor is 0,
1/2, 1 or NaN); db is compared with the FP64 sum (1e-5, FP32) and must be
identical across calls.  The Algodiff op's gradients are checked against
torch's FP64 autograd of the same layer."""
import numpy as np
import pytest

import oracle_ndarray as on
from owl_amd import algodiff as ad
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")

DEV = "cuda:0"

# (x shape, kernel shape, stride, padding, dtypes)
CASES = [
    ([2, 12, 56, 64], [3, 3, 64, 128], [1, 1], SAME, ("bf16", "f16")),   # halo kernel, 128-wide
    ([2, 9, 20, 32], [3, 3, 32, 64], [1, 1], VALID, ("bf16",)),          # halo kernel, 64-wide
    ([2, 10, 10, 16], [3, 3, 16, 32], [1, 1], SAME, ("f32", "bf16")),    # direct-A kernel
    ([2, 11, 9, 5], [3, 2, 5, 7], [2, 1], SAME, ("f32", "f64")),         # generic (unaligned, strided)
    ([1, 5, 6, 4, 8], [2, 3, 2, 8, 16], [1, 2, 1], VALID, ("f32",)),     # 3-D
]
TD = {"f32": torch.float32, "f64": torch.float64, "bf16": torch.bfloat16, "f16": torch.float16}


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dev(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(dt)


def host(t):
    return t.double().cpu().numpy()


def oracle_conv(x, w, stride, pad):
    if x.ndim == 5:
        return on.conv3d(x, w, stride, pad)
    return on.conv2d(x, w, stride, pad)


def fused(x, w, b, stride, pad, **kw):
    f = nd.conv3d_bias_act if x.dim() == 5 else nd.conv2d_bias_act
    return f(x, w, b, stride, pad, **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_fused_forward_vs_oracle(case):
    xs, ks, stride, pad, dts = CASES[case]
    rng = np.random.default_rng(40 + case)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) / np.sqrt(np.prod(ks[:-1]))
    bias = rng.standard_normal(ks[-1]) * 0.5
    for name in dts:
        td = TD[name]
        acc = torch.float64 if name == "f64" else torch.float32
        rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
        b_acc = torch.from_numpy(bias).to(acc)
        ref_pre = oracle_conv(rnd(x), rnd(w), stride, pad) + b_acc.double().numpy()
        ref_y = np.fmax(ref_pre, 0.0)
        out_dt = None if name in ("f32", "f64") else torch.float32
        y, pre = fused(dev(x, td), dev(w, td), b_acc.to(DEV), stride, pad, keep_pre=True, out_dtype=out_dt)
        tol = 1e-12 if name == "f64" else 1e-5
        assert rel(host(pre), ref_pre) < tol, (name, rel(host(pre), ref_pre))
        assert rel(host(y), ref_y) < tol, (name, rel(host(y), ref_y))
        # y is the activation of the stored pre-activation, value for value
        assert torch.equal(y, torch.clamp_min(pre, 0.0)), name
        if name in ("bf16", "f16"):
            # storage-dtype output end to end against the unrounded FP64 oracle
            y16 = fused(dev(x, td), dev(w, td), b_acc.to(DEV), stride, pad)
            ref = np.fmax(oracle_conv(x, w, stride, pad) + bias, 0.0)
            assert y16.dtype == td and rel(host(y16), ref) < 1e-2
        if name == "f32":
            plain = (nd.conv3d if x.ndim == 5 else nd.conv2d)(dev(x, td), dev(w, td), stride, pad)
            unfused = torch.fmax(plain + b_acc.to(DEV), torch.zeros((), device=DEV))
            assert torch.equal(y, unfused)


@pytest.mark.gpu
def test_fused_forward_variants_and_nan():
    rng = np.random.default_rng(7)
    x = rng.standard_normal([1, 8, 8, 32]).astype(np.float32)
    x[0, 3, 4, 5] = np.nan
    w = (rng.standard_normal([3, 3, 32, 64]) / 17).astype(np.float32)
    b = rng.standard_normal(64).astype(np.float32)
    for td in (torch.float32, torch.bfloat16):
        xt, wt = dev(x, td), dev(w, td)
        out_dt = torch.float32
        plain = nd.conv2d(xt, wt, [1, 1], SAME, out_dtype=out_dt)
        bt = dev(b, torch.float32)
        # bias only, relu only, neither
        y_b = nd.conv2d_bias_act(xt, wt, bt, [1, 1], SAME, act="none", out_dtype=out_dt)
        y_r = nd.conv2d_bias_act(xt, wt, None, [1, 1], SAME, act="relu", out_dtype=out_dt)
        y_n = nd.conv2d_bias_act(xt, wt, None, [1, 1], SAME, act=None, out_dtype=out_dt)
        assert torch.equal(torch.nan_to_num(y_b, nan=7.0), torch.nan_to_num(plain + bt, nan=7.0))
        assert torch.equal(y_r, torch.fmax(plain, torch.zeros((), device=DEV)))
        assert torch.equal(torch.nan_to_num(y_n, nan=7.0), torch.nan_to_num(plain, nan=7.0))
        # NaN pre-activations map to 0 through relu, as fmax(NaN, 0) does
        y, pre = nd.conv2d_bias_act(xt, wt, bt, [1, 1], SAME, keep_pre=True, out_dtype=out_dt)
        nan = torch.isnan(pre)
        assert nan.any() and not torch.isnan(y).any() and torch.all(y[nan] == 0)
    with pytest.raises(nd.InvalidArgument):
        nd.conv2d_bias_act(dev(x, torch.float32), dev(w, torch.float32), dev(b[:3], torch.float32), [1, 1])
    with pytest.raises(nd.InvalidArgument):
        nd.conv2d_bias_act(dev(x, torch.float32), dev(w, torch.float32), None, [1, 1], act="gelu")


def _owl_relu_adjoint(pre, dy):
    s = np.where(pre > 0, 1.0, np.where(pre < 0, -1.0, pre))  # signum keeping +-0 and NaN
    return dy * ((s + 1.0) / 2.0)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [[4, 7, 9, 64], [3, 5, 13], [2, 6, 6, 3], [1, 1000, 128]])
@pytest.mark.parametrize("dt", ["f32", "f64", "bf16", "f16"])
def test_bias_act_backward(shape, dt):
    td = TD[dt]
    rng = np.random.default_rng(sum(shape))
    pre = rng.standard_normal(shape)
    pre.reshape(-1)[::17] = 0.0           # exact zeros: Owl's factor 1/2
    pre.reshape(-1)[5::29] = -0.0
    pre.reshape(-1)[3] = np.nan
    dy = rng.standard_normal(shape)
    pt, gt = dev(pre, td), dev(dy, td)
    dz, db = nd.bias_act_backward(pt, gt, act="relu")
    ref = _owl_relu_adjoint(host(pt), host(gt))
    if dt == "f16":
        ref = ref.astype(np.float16).astype(np.float64)   # dy / 2 can land in FP16's subnormals
    got = host(dz)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = ~np.isnan(ref)
    assert np.array_equal(got[fin], ref[fin])            # exact: the factor is 0, 1/2 or 1
    # and the sign of zero: -0 for a negative gradient times the factor 0
    assert np.array_equal(np.signbit(got[fin]), np.signbit(ref[fin]))
    C = shape[-1]
    ref_db = np.nansum(got.reshape(-1, C), axis=0)
    nan_cols = np.isnan(got.reshape(-1, C)).any(axis=0)
    hdb = db.double().cpu().numpy()
    assert np.array_equal(np.isnan(hdb), nan_cols)
    tol = 1e-12 if dt == "f64" else 1e-5
    assert rel(hdb[~nan_cols], ref_db[~nan_cols]) < tol
    # deterministic: the same bits on a second call; no activation copies dy
    _, db2 = nd.bias_act_backward(pt, gt, act="relu")
    assert torch.equal(torch.nan_to_num(db, nan=1.0), torch.nan_to_num(db2, nan=1.0))
    dz0, db0 = nd.bias_act_backward(None, gt, act=None)
    assert torch.equal(dz0, gt)
    assert rel(db0.double().cpu().numpy(), host(gt).reshape(-1, C).sum(axis=0)) < tol


@pytest.mark.gpu
def test_algodiff_conv_bias_relu_vs_torch_autograd():
    rng = np.random.default_rng(3)
    x = rng.standard_normal([2, 9, 8, 6])
    w = rng.standard_normal([3, 3, 6, 5]) / 5
    b = rng.standard_normal(5) * 0.3
    xt, wt, bt = (dev(a, torch.float64) for a in (x, w, b))
    gx, gw, gb = ad.grad(lambda a, k, bb: ad.sum_sqr(ad.conv2d_bias_relu(a, k, bb, [1, 1], SAME)), xt, wt, bt)
    # the same layer in torch (NCHW), FP64 autograd
    X = torch.tensor(x.transpose(0, 3, 1, 2).copy(), requires_grad=True)
    Wt = torch.tensor(w.transpose(3, 2, 0, 1).copy(), requires_grad=True)
    Bt = torch.tensor(b.copy(), requires_grad=True)
    y = torch.relu(torch.nn.functional.conv2d(X, Wt, Bt, padding=1))
    (y * y).sum().backward()
    assert rel(host(gx), X.grad.numpy().transpose(0, 2, 3, 1)) < 1e-12
    assert rel(host(gw), Wt.grad.numpy().transpose(2, 3, 1, 0)) < 1e-12
    assert rel(host(gb), Bt.grad.numpy()) < 1e-12
