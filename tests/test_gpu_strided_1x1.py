"""Backward input of 1x1 convolutions with strides (ResNet's downsampling
layers, Owl's conv2d / conv3d backward input with a 1x1 kernel): dx is
dy . w^T on the strided grid and zero off it.  The library computes the
subsampled grid with the stride-1 backward input and writes dx in one
scatter pass.  Checked against the FP64 oracle on the same rounded inputs
(FP32 / FP64 outputs rel-L2 < 1e-5 and 1e-12, 16-bit < 1e-2), off-grid
values exactly zero, and against the implicit GEMM over every dx pixel
(OWL_AMD_STRIDED_1X1=0) within the same tolerances.  Cases: strides 2 and
3, uneven image sizes, an odd channel count (2-byte rows), 3-D, FP64."""
import os

import numpy as np
import pytest

import oracle_ndarray as on
from owl_amd import _lib
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")

DEV = "cuda:0"

# (x shape, kernel shape, strides, padding, dtype)
CASES = [
    ([4, 28, 28, 64], [1, 1, 64, 32], [2, 2], SAME, "bf16"),      # the halo kernel on the 14x14 grid
    ([2, 29, 31, 48], [1, 1, 48, 40], [3, 3], SAME, "f32"),       # uneven: 10 x 11 grid cells
    ([3, 17, 12, 64], [1, 1, 64, 128], [2, 3], VALID, "f16"),
    ([2, 9, 10, 16], [1, 1, 16, 24], [2, 3], SAME, "f64"),
    ([2, 16, 16, 3], [1, 1, 3, 8], [2, 2], SAME, "bf16"),          # 6-byte dx rows: 2-byte units
    ([2, 8, 9, 10, 32], [1, 1, 1, 32, 16], [2, 2, 2], SAME, "bf16"),  # 3-D
    ([1, 7, 6, 9, 24], [1, 1, 1, 24, 40], [1, 2, 3], VALID, "f32"),   # 3-D, mixed strides
]
TD = {"bf16": torch.bfloat16, "f16": torch.float16, "f32": torch.float32, "f64": torch.float64}
TOL = {"bf16": 1e-2, "f16": 1e-2, "f32": 1e-5, "f64": 1e-12}


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _bwd_input(x, w, stride, dy, generic):
    old = os.environ.get("OWL_AMD_STRIDED_1X1")
    if generic:
        os.environ["OWL_AMD_STRIDED_1X1"] = "0"
    try:
        f = nd.conv2d_backward_input if x.dim() == 4 else nd.conv3d_backward_input
        dx = f(x, w, stride, dy)
        torch.cuda.synchronize()
    finally:
        if old is None:
            os.environ.pop("OWL_AMD_STRIDED_1X1", None)
        else:
            os.environ["OWL_AMD_STRIDED_1X1"] = old
    return dx.double().cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_strided_1x1_bwd_input(case):
    xs, ks, stride, pad, dt = CASES[case]
    print(f"strided 1x1 case {case} x{xs} w{ks} s{stride} {dt}", flush=True)
    td = TD[dt]
    rng = np.random.default_rng(700 + case)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) / np.sqrt(ks[-2])
    ys = (on.conv2d(x, w, stride, pad)).shape
    g = rng.standard_normal(ys)
    rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
    xr, wr, gr = rnd(x), rnd(w), rnd(g)
    ref = on.conv2d_backward_input(xr, wr, stride, gr)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(td)
    xt, wt, gt = dev(x), dev(w), dev(g)
    _lib.kernel_timing(True)
    dx = _bwd_input(xt, wt, stride, gt, generic=False)
    aux = _lib.kernel_time(_lib.KT_AUX)[2]
    _lib.kernel_timing(False)
    assert aux == "stride_scatter_kernel", aux
    assert dx.shape == ref.shape
    assert rel(dx, ref) < TOL[dt], rel(dx, ref)
    # off the strided grid dx is exactly zero
    grid = np.zeros(dx.shape, bool)
    sl = (slice(None),) + tuple(slice(None, None, s) for s in stride) + (slice(None),)
    grid[sl] = True
    assert not np.any(dx[~grid]), np.abs(dx[~grid]).max()
    # the implicit GEMM over every dx pixel agrees
    dg = _bwd_input(xt, wt, stride, gt, generic=True)
    assert rel(dx, dg) < TOL[dt], rel(dx, dg)
    # the same call twice gives the same bits
    assert np.array_equal(dx, _bwd_input(xt, wt, stride, gt, generic=False))
