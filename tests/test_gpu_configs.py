"""The benchmark configurations at their full sizes.

The FP64 oracle is too slow for whole layers, so the results of the full-size
calls (the calls and split-K plans the benchmark times) are checked by:
  * forward and backward-input on the first two images and on two images
    drawn from the rest of the batch (each output image depends on its own
    input image only), against the oracle;
  * backward-kernel on 256 sampled entries, each an FP64 dot product over all
    B * Ho * Wo pixels computed on the host (``fullsize.dw_entries``), plus
    the whole 2-image slice run on its own (a different split plan);
  * the adjoint identity <conv(x,w), g> = <x, dgrad(w,g)> = <w, wgrad(x,g)>
    evaluated in FP64 on the host from the GPU's outputs, within tol times
    ||y * g||_2 (the scale of the rounding error of the random sum), and only
    where <y, g> is large enough that an all-zero result would fail it.
Tolerances: FP32 rel-L2 1e-5 against FP64; BF16 / FP16 outputs 1e-2; the
16-bit layers' FP32 weight gradient 1e-5 against FP64 on the rounded inputs."""
import numpy as np
import pytest

import oracle_ndarray as on
from fullsize import dw_entries, pick_images, sample_dw_index, sampled_rel
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")
DEV = "cuda:0"


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def layer(xs, ks, stride, dtype, seed, uniform=False):
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    if uniform:
        x = torch.rand(xs, device=DEV, generator=g) * 2 - 1
        w = torch.rand(ks, device=DEV, generator=g) * 2 - 1
    else:
        x = torch.randn(xs, device=DEV, generator=g)
        fan = int(np.prod(ks[:-1]))
        w = torch.randn(ks, device=DEV, generator=g) * (2.0 / fan) ** 0.5
    return x.to(dtype).contiguous(), w.to(dtype).contiguous(), g


def h64(t):
    return t.double().cpu().numpy()


@pytest.mark.gpu
def test_config1_conv2d_forward_fp32():
    """[8,32,32,3] (*) [3,3,3,16] SAME, uniform[-1,1]: whole output vs FP64."""
    x, w, _ = layer([8, 32, 32, 3], [3, 3, 3, 16], [1, 1], torch.float32, 1, uniform=True)
    y = nd.conv2d(x, w, [1, 1], SAME)
    r = on.conv2d(h64(x), h64(w), [1, 1], SAME)
    assert tuple(y.shape) == (8, 32, 32, 16)
    assert rel(h64(y), r) < 1e-5
    assert np.abs(h64(y) - r).max() < 1e-5 * np.abs(r).max() * 10


def check_full_layer(x, w, stride, dy, y, dx, dw, tol, seed):
    """The whole-layer results y, dx (x.dtype) and dw (FP32) of one call each."""
    n = 2
    ws = h64(w)
    for im in ([0, 1], pick_images(x.shape[0], n, seed)):
        xs, gs = h64(x[im]), h64(dy[im])
        assert rel(h64(y[im]), on.conv2d(xs, ws, stride, SAME)) < tol, im
        assert rel(h64(dx[im]), on.conv2d_backward_input(xs, ws, stride, gs)) < tol, im
    # the full-size weight gradient: 256 entries against FP64 dot products
    # over every pixel (the rounded inputs: the 16-bit passes accumulate and
    # store dw in FP32)
    xf, gf = x.float().cpu().numpy(), dy.float().cpu().numpy()
    idx = sample_dw_index(list(w.shape[:2]), w.shape[2], w.shape[3], 256, seed)
    ref = dw_entries(xf, gf, list(w.shape[:2]), stride, idx)
    got = dw.cpu().numpy()[tuple(idx.T)]
    assert sampled_rel(got, ref) < 1e-5, sampled_rel(got, ref)
    # the 2-image slice on its own (a different split plan)
    dws = nd.conv2d_backward_kernel(x[:n].contiguous(), w, stride, dy[:n].contiguous(), out_dtype=torch.float32)
    assert rel(h64(dws), on.conv2d_backward_kernel(h64(x[:n]), ws, stride, h64(dy[:n]))) < max(tol, 1e-5)
    # adjoint identity at full size, normalised by ||y * dy||_2
    yd, dyd = h64(y), h64(dy)
    lhs = float((yd * dyd).sum())
    bound = max(tol, 1e-5) * float(np.linalg.norm(yd * dyd))
    assert abs(lhs) > 3 * bound, "uninformative adjoint check: choose another seed"
    assert abs(lhs - float((h64(x) * h64(dx)).sum())) < bound
    assert abs(lhs - float((h64(w) * h64(dw)).sum())) < bound


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.bfloat16, 1e-2)], ids=["fp32", "bf16"])
def test_config2_vgg_layer(dtype, tol):
    x, w, g = layer([64, 56, 56, 64], [3, 3, 64, 128], [1, 1], dtype, 2)
    dy = torch.randn([64, 56, 56, 128], device=DEV, generator=g).to(dtype)
    y = nd.conv2d(x, w, [1, 1], SAME)
    dx = nd.conv2d_backward_input(x, w, [1, 1], dy)
    dw = nd.conv2d_backward_kernel(x, w, [1, 1], dy, out_dtype=torch.float32)
    torch.cuda.synchronize()
    check_full_layer(x, w, [1, 1], dy, y, dx, dw, tol, 20)


@pytest.mark.gpu
def test_config2_maxpool_bit_exact():
    x, w, g = layer([64, 56, 56, 64], [3, 3, 64, 128], [1, 1], torch.float32, 3)
    y = nd.conv2d(x, w, [1, 1], SAME)
    p, a = nd.max_pool2d_argmax(y, [2, 2], [2, 2], VALID)
    gp = torch.randn(tuple(p.shape), device=DEV, generator=g)
    dx = nd.max_pool2d_backward(VALID, y, [2, 2], [2, 2], gp)
    yh = y.cpu().numpy()
    rp, ra = on.max_pool2d_argmax(yh, [2, 2], [2, 2], VALID)
    assert np.array_equal(p.cpu().numpy(), rp)
    assert np.array_equal(a.cpu().numpy(), ra)
    assert np.array_equal(dx.cpu().numpy(), on.max_pool2d_backward(VALID, yh, [2, 2], [2, 2], gp.cpu().numpy()))


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.bfloat16, 1e-2), (torch.float16, 1e-2)],
                         ids=["fp32", "bf16", "fp16"])
def test_config3_resnet_stack(dtype, tol):
    """[128,28,28,256]: 3x3 256->256 stride 1 SAME, then 1x1 256->512 stride 2."""
    x, w, g = layer([128, 28, 28, 256], [3, 3, 256, 256], [1, 1], dtype, 4)
    w2 = (torch.randn([1, 1, 256, 512], device=DEV, generator=g) * (2.0 / 256) ** 0.5).to(dtype)
    for i, (ww, stride) in enumerate(((w, [1, 1]), (w2, [2, 2]))):
        y = nd.conv2d(x, ww, stride, SAME)
        dy = torch.randn(tuple(y.shape), device=DEV, generator=g).to(dtype)
        dx = nd.conv2d_backward_input(x, ww, stride, dy)
        dw = nd.conv2d_backward_kernel(x, ww, stride, dy, out_dtype=torch.float32)
        torch.cuda.synchronize()
        check_full_layer(x, ww, stride, dy, y, dx, dw, tol, 30 + i)


@pytest.mark.gpu
def test_config3_resnet_stack_fp8_forward():
    """Config 3's FP8 e4m3 forward at full size: per-tensor amax scales, the
    3x3 256->256 layer (FP8 halo kernel) and the 1x1 256->512 stride-2 layer
    (generic kernel), FP32 and BF16 output.  On a 2-image slice: within 1e-4
    (FP32 output; 1e-2 for BF16) of the FP64 oracle on the dequantised
    operands, the MFMA's accumulation being the only error, and rel-L2 < 8e-2
    against the FP64 oracle on the unquantised inputs (e4m3 keeps 3 mantissa
    bits)."""
    import ctypes
    from owl_amd import _lib
    x, w, g = layer([128, 28, 28, 256], [3, 3, 256, 256], [1, 1], torch.float32, 6)
    w2 = torch.randn([1, 1, 256, 512], device=DEV, generator=g) * (2.0 / 256) ** 0.5
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def quant(t):
        q = torch.empty(t.numel(), dtype=torch.uint8, device=DEV)
        sc = torch.empty(1, dtype=torch.float32, device=DEV)
        _lib.check(_lib.lib().owl_amd_fp8_quantize(_lib.F32, ctypes.c_void_p(t.data_ptr()), t.numel(),
                                                   ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(sc.data_ptr()), s))
        return q.view(t.shape), sc

    def deq(q, sc):
        return (q.view(torch.float8_e4m3fn).float().double() * float(sc.item())).cpu().numpy()

    xq, sx = quant(x)
    n = 2
    for ww, stride in ((w, [1, 1]), (w2, [2, 2])):
        wq, sw = quant(ww)
        geo = nd.conv_problem(list(x.shape), list(ww.shape), stride, None, SAME)
        d = geo.desc()
        rdq = on.conv2d(deq(xq[:n].contiguous(), sx), deq(wq, sw), stride, SAME)
        ry = on.conv2d(h64(x[:n]), h64(ww), stride, SAME)
        for td, ot, tol in ((torch.float32, _lib.F32, 1e-4), (torch.bfloat16, _lib.BF16, 1e-2)):
            y = torch.empty([128] + geo.sp_out + [ww.shape[3]], dtype=td, device=DEV)
            _lib.check(_lib.lib().owl_amd_conv_fwd_fp8(ctypes.byref(d), ot, ctypes.c_void_p(xq.data_ptr()),
                                                       ctypes.c_void_p(wq.data_ptr()), ctypes.c_void_p(sx.data_ptr()),
                                                       ctypes.c_void_p(sw.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                                       s))
            torch.cuda.synchronize()
            assert np.isfinite(h64(y)).all()
            ys = h64(y[:n])
            assert rel(ys, rdq) < tol, (stride, td, rel(ys, rdq))
            assert rel(ys, ry) < 8e-2, (stride, td, rel(ys, ry))


@pytest.mark.gpu
def test_config5_conv3d_stack_fp32():
    """conv3d [8,32,32,32,16] (*) [3,3,3,16,32] SAME then (*) [3,3,3,32,64], avg_pool3d 2x2x2."""
    x, w1, g = layer([8, 32, 32, 32, 16], [3, 3, 3, 16, 32], [1, 1, 1], torch.float32, 5)
    w2 = (torch.randn([3, 3, 3, 32, 64], device=DEV, generator=g) * (2.0 / 864) ** 0.5)
    y1 = nd.conv3d(x, w1, [1, 1, 1], SAME)
    y2 = nd.conv3d(y1, w2, [1, 1, 1], SAME)
    p = nd.avg_pool3d(y2, [2, 2, 2], [2, 2, 2], VALID)
    n = 1
    r1 = on.conv3d(h64(x[:n]), h64(w1), [1, 1, 1], SAME)
    assert rel(h64(y1[:n]), r1) < 1e-5
    r2 = on.conv3d(h64(y1[:n]), h64(w2), [1, 1, 1], SAME)
    assert rel(h64(y2[:n]), r2) < 1e-5
    y2h = y2[:n].cpu().numpy()
    assert np.array_equal(p[:n].cpu().numpy(), on.avg_pool3d(y2h, [2, 2, 2], [2, 2, 2], VALID))
