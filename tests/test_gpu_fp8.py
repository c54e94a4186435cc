"""FP8 (OCP e4m3) forward: the encoder against torch's float8_e4m3fn cast,
the per-tensor quantiser, and the FP8-MFMA conv against the FP64 oracle.

Tolerances: on the dequantised inputs (q * scale) the FP8 conv adds only the
MFMA's accumulation error; the f8f6f4 MFMA sums its 64 products per step less
exactly than an FP32 FMA chain (measured ~1.4e-5 rel-L2) -> bound 1e-4.  Against the unquantised FP64 oracle the
e4m3 rounding of both operands (3 mantissa bits, ~2^-4 / sqrt(3) rms per
operand) bounds the error -> rel-L2 < 8e-2 (reported per layer by the
precision sweep, not forced under 1e-2)."""
import ctypes

import numpy as np
import pytest

import oracle_ndarray as on
from owl_amd import _lib
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")
DEV = "cuda:0"


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.gpu
def test_e4m3_encoder_matches_torch():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.standard_normal(20000) * s for s in (1e-3, 1e-1, 1, 10, 100)]).astype(np.float32)
    v = np.concatenate([v, np.array([0.0, -0.0, 448.0, -448.0, 2 ** -9, 2 ** -10, 3 * 2 ** -10, 0.015625,
                                     0.0146484375, 1.0625, 1.1875, 240.0, 232.0, 3.0], np.float32)])
    v = v[np.abs(v) <= 448]
    src = torch.from_numpy(v).to(DEV)
    dst = torch.empty(v.size, dtype=torch.uint8, device=DEV)
    _lib.check(_lib.lib().owl_amd_fp8_encode(ctypes.c_void_p(src.data_ptr()), v.size,
                                             ctypes.c_void_p(dst.data_ptr()), _s()))
    ref = torch.from_numpy(v).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(dst.cpu().numpy(), ref)
    # saturation at +-448 instead of NaN
    big = torch.tensor([500.0, -1e6, 449.0], device=DEV)
    out = torch.empty(3, dtype=torch.uint8, device=DEV)
    _lib.check(_lib.lib().owl_amd_fp8_encode(ctypes.c_void_p(big.data_ptr()), 3, ctypes.c_void_p(out.data_ptr()),
                                             _s()))
    assert out.cpu().numpy().tolist() == [0x7E, 0xFE, 0x7E]


def quantize(t):
    q = torch.empty(t.numel(), dtype=torch.uint8, device=DEV)
    sc = torch.empty(1, dtype=torch.float32, device=DEV)
    dt = _lib.F32 if t.dtype == torch.float32 else _lib.BF16
    _lib.check(_lib.lib().owl_amd_fp8_quantize(dt, ctypes.c_void_p(t.data_ptr()), t.numel(),
                                               ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(sc.data_ptr()), _s()))
    return q.view(t.shape), sc


@pytest.mark.gpu
@pytest.mark.parametrize("dt,n,offset", [(torch.bfloat16, 1000003, 0), (torch.float32, 1000001, 0),
                                         (torch.bfloat16, 4099, 1), (torch.float32, 77, 3)])
def test_quantizer_matches_torch(dt, n, offset):
    """The per-tensor quantiser: scale = amax / 448 and q = e4m3(x * (448 /
    amax)), byte for byte against torch's e4m3 cast of the same FP32
    products; lengths that leave a tail past the last 16-byte chunk, and
    sources offset from 16-byte alignment (the element-by-element path)."""
    g = torch.Generator(device=DEV)
    g.manual_seed(n)
    base = (torch.randn(n + offset, device=DEV, generator=g) * 3).to(dt)
    t = base[offset:]
    q = torch.empty(n, dtype=torch.uint8, device=DEV)
    sc = torch.empty(1, dtype=torch.float32, device=DEV)
    code = _lib.F32 if dt == torch.float32 else _lib.BF16
    _lib.check(_lib.lib().owl_amd_fp8_quantize(code, ctypes.c_void_p(t.data_ptr()), n,
                                               ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(sc.data_ptr()), _s()))
    torch.cuda.synchronize()
    v = t.float()
    amax = v.abs().max()
    inv = torch.tensor(448.0, device=DEV) / amax
    assert sc.item() == (amax / torch.tensor(448.0, device=DEV)).item()
    ref = (v * inv).to(torch.float8_e4m3fn).view(torch.uint8)
    assert torch.equal(q, ref)


def dequant(q, sc):
    f = q.view(torch.float8_e4m3fn).float()
    return (f.double() * float(sc.item())).cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [([2, 9, 9, 32], [3, 3, 32, 48], [1, 1], SAME),
                                   ([3, 8, 10, 64], [1, 1, 64, 32], [2, 2], SAME),
                                   ([2, 7, 7, 5], [3, 3, 5, 7], [1, 2], VALID)])
def test_fp8_conv_forward(shape):
    xs, ks, stride, pad = shape
    rng = np.random.default_rng(3)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) * np.sqrt(2.0 / np.prod(ks[:3]))
    xt = torch.from_numpy(x).float().to(DEV).contiguous()
    wt = torch.from_numpy(w).float().to(DEV).contiguous()
    xq, sx = quantize(xt)
    wq, sw = quantize(wt)
    assert abs(float(sx.item()) - np.abs(xt.cpu().numpy()).max() / 448) < 1e-6 * float(sx.item())
    ry = on.conv2d(x, w, stride, pad)
    y = torch.empty(ry.shape, dtype=torch.float32, device=DEV)
    from owl_amd import ndarray as nd
    g = nd._Conv(xs[0], xs[1:3], xs[3], ks[:2], ks[3], ry.shape[1:3], stride, [1, 1],
                 nd._pads_from_shapes(xs[1:3], ks[:2], ry.shape[1:3], stride, [1, 1]) if pad == SAME else [0, 0])
    d = g.desc()
    _lib.check(_lib.lib().owl_amd_conv_fwd_fp8(ctypes.byref(d), _lib.F32, ctypes.c_void_p(xq.data_ptr()),
                                               ctypes.c_void_p(wq.data_ptr()), ctypes.c_void_p(sx.data_ptr()),
                                               ctypes.c_void_p(sw.data_ptr()), ctypes.c_void_p(y.data_ptr()), _s()))
    torch.cuda.synchronize()
    yg = y.double().cpu().numpy()
    rdq = on.conv2d(dequant(xq, sx), dequant(wq, sw), stride, pad)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(yg, rdq) < 1e-4
    assert rel(yg, ry) < 8e-2


# Layers the FP8 halo kernel takes (2-D, stride 1, Ci % 64 == 0): the VGG
# layer's shape, a second 64-channel slice only partly used, a 1x1 kernel,
# dilation with VALID padding and a short last row tile; BF16 and FP32 output.
HALO_FP8 = [([2, 12, 56, 64], [3, 3, 64, 128], 1, SAME),
            ([2, 9, 20, 128], [3, 3, 128, 72], 1, VALID),
            ([1, 7, 30, 64], [1, 1, 64, 64], 1, SAME),
            ([2, 12, 20, 64], [3, 3, 64, 96], 2, VALID),
            ([3, 13, 56, 64], [3, 3, 64, 128], 1, SAME)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(HALO_FP8)))
@pytest.mark.parametrize("out", ["f32", "bf16"])
def test_fp8_halo_forward(case, out):
    xs, ks, rate, pad = HALO_FP8[case]
    print(f"fp8 halo case {case} out {out} x{xs} w{ks} rate {rate}", flush=True)
    rng = np.random.default_rng(40 + case)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) * np.sqrt(2.0 / np.prod(ks[:3]))
    xt = torch.from_numpy(x).float().to(DEV).contiguous()
    wt = torch.from_numpy(w).float().to(DEV).contiguous()
    xq, sx = quantize(xt)
    wq, sw = quantize(wt)
    from owl_amd import ndarray as nd
    g = nd.conv_problem(xs, ks, [1, 1], [rate, rate], pad)
    ry = on.dilated_conv2d(x, w, [1, 1], [rate, rate], pad)
    assert list(ry.shape[1:3]) == list(g.sp_out)
    td, ot = (torch.float32, _lib.F32) if out == "f32" else (torch.bfloat16, _lib.BF16)
    y = torch.full(ry.shape, float("nan"), dtype=td, device=DEV)
    d = g.desc()
    _lib.check(_lib.lib().owl_amd_conv_fwd_fp8(ctypes.byref(d), ot, ctypes.c_void_p(xq.data_ptr()),
                                               ctypes.c_void_p(wq.data_ptr()), ctypes.c_void_p(sx.data_ptr()),
                                               ctypes.c_void_p(sw.data_ptr()), ctypes.c_void_p(y.data_ptr()), _s()))
    y2 = torch.full(ry.shape, float("nan"), dtype=td, device=DEV)
    _lib.check(_lib.lib().owl_amd_conv_fwd_fp8(ctypes.byref(d), ot, ctypes.c_void_p(xq.data_ptr()),
                                               ctypes.c_void_p(wq.data_ptr()), ctypes.c_void_p(sx.data_ptr()),
                                               ctypes.c_void_p(sw.data_ptr()), ctypes.c_void_p(y2.data_ptr()), _s()))
    torch.cuda.synchronize()
    yg = y.double().cpu().numpy()
    assert np.isfinite(yg).all()
    assert torch.equal(y, y2)  # deterministic, every output written
    rdq = on.dilated_conv2d(dequant(xq, sx), dequant(wq, sw), [1, 1], [rate, rate], pad)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(yg, rdq) < (1e-4 if out == "f32" else 1e-2), rel(yg, rdq)
    assert rel(yg, ry) < 8e-2
