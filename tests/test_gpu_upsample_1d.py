"""Upsampling and the 1-D convolution / pooling family on the GPU.

Upsampling: the reference's own test cases (test/unit_upsampling_generic.ml,
through the host drop-ins and on device tensors) and random data against the
oracle restatement -- bit-exact for FP32 and FP64 (the backward adds in Owl's
order), including output extents past the scaled size, which Owl's stub
clamps onto the last row / column; BF16 within one rounding of the FP32
result.

1-D ops: Owl runs them as the 2-D ops on [B; 1; W; C]; the results are checked
against a direct 1-D convolution / window pooling written out here in numpy
(FP64, 1e-12), and the backward passes against the adjoint identities
<conv(x), g> = <x, bwd_input(g)> = <w, bwd_kernel(g)>."""
import json
import os

import numpy as np
import pytest

import oracle
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID, conv_out_dim

torch = pytest.importorskip("torch")

DEV = "cuda:0"
HERE = os.path.dirname(os.path.abspath(__file__))
UP = json.load(open(os.path.join(HERE, "golden", "owl_upsampling_vectors.json")))["cases"]


def seq(shape, dt):
    return np.arange(int(np.prod(shape)), dtype=dt).reshape(shape)


@pytest.mark.gpu
@pytest.mark.parametrize("case", UP, ids=[c["fun"] for c in UP])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_upsampling_owl_cases(case, dt):
    shp, size = case["input_shape"], case["size"]
    s0, s1 = size
    for on_dev in (False, True):
        conv = (lambda a: torch.from_numpy(a).to(DEV)) if on_dev else (lambda a: a)
        back = (lambda t: t.cpu().numpy()) if on_dev else (lambda a: a)
        if case["helper"] == "test_upsampling2d_forward":
            x = seq(shp, dt)
            y = back(nd.upsampling2d(conv(x), size))
            assert np.array_equal(y, x.repeat(s0, axis=1).repeat(s1, axis=2))
        else:
            dy = seq([shp[0], shp[1] * s0, shp[2] * s1, shp[3]], dt)
            dx = back(nd.upsampling2d_backward(conv(np.ones(shp, dt)), size, conv(dy)))
            assert np.array_equal(dx.ravel(), np.array(case["expected"], dt))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_upsampling_random_vs_oracle(seed):
    rng = np.random.default_rng(seed)
    B, H, W, C = 2, int(rng.integers(1, 9)), int(rng.integers(1, 9)), int(rng.choice([1, 3, 8, 16, 64]))
    s0, s1 = int(rng.integers(1, 4)), int(rng.integers(1, 4))
    for dt in (np.float32, np.float64):
        x = rng.standard_normal([B, H, W, C]).astype(dt)
        g = rng.standard_normal([B, H * s0, W * s1, C]).astype(dt)
        assert np.array_equal(nd.upsampling2d(x, [s0, s1]), oracle.upsampling_fwd(x, [s0, s1]))
        ref = oracle.upsampling_bwd(g, x.shape, [s0, s1])
        assert np.array_equal(nd.upsampling2d_backward(x, [s0, s1], g), ref)
        gd = torch.from_numpy(g).to(DEV)
        assert np.array_equal(nd.upsampling2d_backward(torch.from_numpy(x).to(DEV), [s0, s1], gd).cpu().numpy(), ref)
    # BF16 device tensors: one rounding of the FP32 sums
    xb = torch.from_numpy(x.astype(np.float32)).to(DEV).bfloat16()
    gb = torch.from_numpy(g.astype(np.float32)).to(DEV).bfloat16()
    yb = nd.upsampling2d(xb, [s0, s1])
    assert torch.equal(yb, xb.repeat_interleave(s0, 1).repeat_interleave(s1, 2))
    db = nd.upsampling2d_backward(xb, [s0, s1], gb).float().cpu().numpy()
    ref32 = oracle.upsampling_bwd(gb.float().cpu().numpy(), x.shape, [s0, s1])
    assert np.allclose(db, ref32, rtol=2 ** -8, atol=1e-6)


@pytest.mark.gpu
def test_upsampling_clamped_extent_matches_owl_stub():
    # the stub takes explicit output extents; cells past H * s0 / W * s1 go to
    # the last input row / column (owl_ndarray_upsampling_impl.h:41-48)
    rng = np.random.default_rng(9)
    for dt in (np.float32, np.float64):
        dy = rng.standard_normal([2, 7, 5, 3]).astype(dt)
        dx = np.empty([2, 3, 2, 3], dt)
        nd._lib.call(f"owl_amd_{'float32' if dt == np.float32 else 'float64'}_ndarray_upsampling_spatial_backward",
                     nd._ptr(dx), nd._ptr(dy), 2, 3, 2, 3, 7, 5, 2, 2)
        assert np.array_equal(dx, oracle.upsampling_bwd(dy, [2, 3, 2, 3], [2, 2]))
    with pytest.raises(nd.InvalidArgument):
        nd.upsampling2d(np.ones([2, 3, 3], np.float32), [2, 2])
    with pytest.raises(nd.InvalidArgument):
        nd.upsampling2d_backward(np.ones([1, 2, 2, 1], np.float32), [2, 2], np.ones([1, 4, 5, 1], np.float32))


# ---------------------------------------------------------------- 1-D family
def direct_conv1d(x, w, s, padding):
    B, W, Ci = x.shape
    kw, _, Co = w.shape
    Wo = conv_out_dim(padding, W, kw, s)
    pad = max(0, ((Wo - 1) * s + kw - W) // 2) if padding == SAME else 0
    y = np.zeros([B, Wo, Co])
    for j in range(Wo):
        for c in range(kw):
            src = j * s + c - pad
            if 0 <= src < W:
                y[:, j, :] += x[:, src, :] @ w[c]
    return y


def direct_pool1d(x, k, s, padding, is_max):
    B, W, C = x.shape
    Wo = conv_out_dim(padding, W, k, s)
    pad = max(0, ((Wo - 1) * s + k - W) // 2) if padding == SAME else 0
    y = np.empty([B, Wo, C])
    for j in range(Wo):
        lo, hi = max(0, j * s - pad), min(W, j * s - pad + k)
        win = x[:, lo:hi, :]
        y[:, j, :] = win.max(axis=1) if is_max else win.mean(axis=1)
    return y


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(6))
def test_conv1d_family(seed):
    rng = np.random.default_rng(seed)
    B, W, Ci, Co = 2, int(rng.integers(5, 30)), int(rng.choice([1, 3, 8])), int(rng.choice([2, 5, 16]))
    kw, s = int(rng.integers(1, 5)), int(rng.integers(1, 3))
    padding = [SAME, VALID][seed % 2]
    x = rng.standard_normal([B, W, Ci])
    w = rng.standard_normal([kw, Ci, Co])
    for on_dev in (False, True):
        conv = (lambda a: torch.from_numpy(a).to(DEV)) if on_dev else (lambda a: a)
        back = (lambda t: t.cpu().numpy()) if on_dev else (lambda a: a)
        y = back(nd.conv1d(conv(x), conv(w), [s], padding))
        ref = direct_conv1d(x, w, s, padding)
        assert y.shape == ref.shape and np.abs(y - ref).max() < 1e-12 * max(1, np.abs(ref).max())
        g = rng.standard_normal(ref.shape)
        dx = back(nd.conv1d_backward_input(conv(x), conv(w), [s], conv(g)))
        dw = back(nd.conv1d_backward_kernel(conv(x), conv(w), [s], conv(g)))
        ip = np.vdot(ref, g)
        assert abs(np.vdot(x, dx) - ip) < 1e-10 * max(1, abs(ip))
        assert abs(np.vdot(w, dw) - ip) < 1e-10 * max(1, abs(ip))
    # dilation: the 2-D rate is passed through, as in Owl
    r = [1, 2]
    yd = nd.dilated_conv1d(x, w, [1], r, VALID)
    we = np.zeros([(kw - 1) * 2 + 1, Ci, Co])
    we[::2] = w
    assert np.abs(yd - direct_conv1d(x, we, 1, VALID)).max() < 1e-12 * max(1, np.abs(yd).max())
    with pytest.raises(nd.InvalidArgument):
        nd.dilated_conv1d(x, w, [1], [2], VALID)     # a 1-entry rate fails in the 2-D check, as in Owl
    # transpose: adjoint of the forward conv with the kernel read as [kw; co; ci]
    wt = rng.standard_normal([kw, Co, Ci])
    xt = rng.standard_normal([B, W, Co])
    yt = nd.transpose_conv1d(xt, wt, [s], padding)
    assert yt.shape[1] == (W * s if padding == SAME else W * s + max(kw - s, 0))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(6))
def test_pool1d_family(seed):
    rng = np.random.default_rng(100 + seed)
    B, W, C = 2, int(rng.integers(4, 25)), int(rng.choice([1, 4, 16]))
    k, s = int(rng.integers(1, 4)), int(rng.integers(1, 4))
    padding = [SAME, VALID][seed % 2]
    if conv_out_dim(padding, W, k, s) <= 0:
        return
    x = rng.standard_normal([B, W, C])
    for on_dev in (False, True):
        conv = (lambda a: torch.from_numpy(a).to(DEV)) if on_dev else (lambda a: a)
        back = (lambda t: t.cpu().numpy()) if on_dev else (lambda a: a)
        assert np.array_equal(back(nd.max_pool1d(conv(x), [k], [s], padding)), direct_pool1d(x, k, s, padding, True))
        ya = back(nd.avg_pool1d(conv(x), [k], [s], padding))
        assert np.abs(ya - direct_pool1d(x, k, s, padding, False)).max() < 1e-14
        g = rng.standard_normal(ya.shape)
        dmax = back(nd.max_pool1d_backward(padding, conv(x), [k], [s], conv(g)))
        davg = back(nd.avg_pool1d_backward(padding, conv(x), [k], [s], conv(g)))
        assert abs(dmax.sum() - g.sum()) < 1e-10 and abs(davg.sum() - g.sum()) < 1e-10  # mass conservation
        x2 = x.reshape(B, 1, W, C)
        assert np.array_equal(dmax, nd.max_pool2d_backward(padding, x2, [1, k], [1, s],
                                                           g.reshape(B, 1, -1, C)).reshape(x.shape))
