"""Reverse-mode gradients through the GPU conv3d / avg_pool3d custom ops
(the 3-D differentiation workload): conv3d [8,32,32,32,16] (*) [3,3,3,16,32]
then (*) [3,3,3,32,64], stride 1 SAME, avg_pool3d 2x2x2, loss = sum of
squares.  Gradients w.r.t. the input and both kernels, in FP32 and BF16, are
checked against FP64 central finite differences (eps 1e-4) at 256 random
coordinates, and (on a 1-image batch) against the FP64 oracle's chain rule
through Owl's conv3d_backward_* / avg_pool3d_backward."""
import numpy as np
import pytest

import oracle_ndarray as on
from owl_amd import algodiff as ad
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")
DEV = "cuda:0"


def params(B, seed):
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    x = torch.randn([B, 32, 32, 32, 16], device=DEV, generator=g, dtype=torch.float64)
    w1 = torch.randn([3, 3, 3, 16, 32], device=DEV, generator=g, dtype=torch.float64) * (2 / 432) ** 0.5
    w2 = torch.randn([3, 3, 3, 32, 64], device=DEV, generator=g, dtype=torch.float64) * (2 / 864) ** 0.5
    return x, w1, w2


def net(x, w1, w2):
    y = ad.conv3d(ad.conv3d(x, w1, [1, 1, 1], SAME), w2, [1, 1, 1], SAME)
    return ad.sum_sqr(ad.avg_pool3d(y, [2, 2, 2], [2, 2, 2], VALID))


def loss64(x, w1, w2):
    return net(ad.make_reverse(x), ad.make_reverse(w1), ad.make_reverse(w2)).primal


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 2e-3), (torch.bfloat16, 6e-2)], ids=["fp32", "bf16"])
def test_conv3d_stack_gradients_vs_finite_differences(dtype, tol):
    x, w1, w2 = params(8, 11)
    gx, gw1, gw2 = ad.grad(net, x.to(dtype), w1.to(dtype), w2.to(dtype))
    grads = [gx.double(), gw1.double(), gw2.double()]
    rng = np.random.default_rng(5)
    p64 = [x, w1, w2]
    errs, scale = [], []
    for _ in range(256):
        k = int(rng.integers(3))
        idx = tuple(int(rng.integers(s)) for s in p64[k].shape)
        fd = ad.finite_difference(loss64, p64, (k, *idx), eps=1e-4)
        errs.append(abs(float(grads[k][idx]) - fd))
        scale.append(abs(fd))
    # normalised like Owl_algodiff_check: max error over the RMS of the FD gradient
    rms = float(np.sqrt(np.mean(np.square(scale))))
    assert max(errs) / rms < tol, (max(errs), rms)


@pytest.mark.gpu
def test_conv3d_stack_gradients_vs_oracle_chain():
    x, w1, w2 = params(1, 12)
    gx, gw1, gw2 = ad.grad(net, x.float(), w1.float(), w2.float())
    xh, w1h, w2h = (t.cpu().numpy() for t in (x.float().double(), w1.float().double(), w2.float().double()))
    y1 = on.conv3d(xh, w1h, [1, 1, 1], SAME)
    y2 = on.conv3d(y1, w2h, [1, 1, 1], SAME)
    p = on.avg_pool3d(y2, [2, 2, 2], [2, 2, 2], VALID)
    dy2 = on.avg_pool3d_backward(VALID, y2, [2, 2, 2], [2, 2, 2], 2.0 * p)
    dw2 = on.conv3d_backward_kernel(y1, w2h, [1, 1, 1], dy2)
    dy1 = on.conv3d_backward_input(y1, w2h, [1, 1, 1], dy2)
    dw1 = on.conv3d_backward_kernel(xh, w1h, [1, 1, 1], dy1)
    dx = on.conv3d_backward_input(xh, w1h, [1, 1, 1], dy1)
    rel = lambda a, b: float(np.linalg.norm(a.double().cpu().numpy() - b) / np.linalg.norm(b))
    assert rel(gx, dx) < 1e-5 and rel(gw1, dw1) < 1e-5 and rel(gw2, dw2) < 1e-5
