"""Backward input of strided convolutions with kernels larger than 1x1
(ResNet's 3x3 stride-2 layers, and the forward of Owl's transpose
convolutions): the library splits dx into its stride phases, runs each as a
stride-1 backward input with the kernel taps of that phase, and interleaves
the phase grids into dx (phases without taps are zeros).  Checked against
the FP64 oracle on the same rounded inputs (FP32 / FP64 outputs rel-L2 <
1e-5 and 1e-12, 16-bit < 1e-2) and against the implicit GEMM over every dx
pixel (OWL_AMD_STRIDED_PHASES=0).  Cases: kernels 2, 3 and 5 with strides 2
and 3 (some phases without taps when the kernel is shorter than the
stride), SAME and VALID, uneven sizes, mixed strides, 3-D, every real
dtype.  Also the forward of strided convolutions whose dilation is a
multiple of the stride, which read one stride phase of x, and strided 2-D
forwards as a stride-1 convolution of x's phases stacked as channels
(space-to-depth): against the oracle and the direct paths
(OWL_AMD_STRIDED_FWD_PHASE=0, OWL_AMD_S2D=0)."""
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
    ([4, 28, 28, 64], [3, 3, 64, 64], [2, 2], SAME, "bf16"),       # ResNet-style 3x3 stride 2
    ([2, 29, 31, 32], [3, 3, 32, 48], [2, 2], SAME, "f32"),       # uneven image
    ([3, 17, 12, 40], [5, 5, 40, 32], [2, 3], SAME, "f16"),       # 5x5, mixed strides
    ([2, 13, 14, 16], [3, 2, 16, 24], [3, 3], VALID, "f64"),      # kernel shorter than the stride: empty phases
    ([2, 16, 16, 8], [2, 2, 8, 16], [2, 2], VALID, "bf16"),       # 2x2 stride 2: one tap per phase
    ([2, 15, 15, 24], [3, 3, 24, 32], [3, 3], SAME, "bf16"),      # stride 3
    ([2, 8, 9, 10, 32], [3, 3, 3, 32, 16], [2, 2, 2], SAME, "bf16"),  # 3-D
    ([1, 7, 6, 9, 16], [2, 3, 3, 16, 24], [1, 2, 3], VALID, "f32"),   # 3-D, mixed strides
]
TD = {"bf16": torch.bfloat16, "f16": torch.float16, "f32": torch.float32, "f64": torch.float64}
TOL = {"bf16": 1e-2, "f16": 1e-2, "f32": 1e-5, "f64": 1e-12}


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _bwd_input(x, w, stride, dy, generic):
    old = os.environ.get("OWL_AMD_STRIDED_PHASES")
    if generic:
        os.environ["OWL_AMD_STRIDED_PHASES"] = "0"
    try:
        f = nd.conv2d_backward_input if x.dim() == 4 else nd.conv3d_backward_input
        dx = f(x, w, stride, dy)
        torch.cuda.synchronize()
    finally:
        if old is None:
            os.environ.pop("OWL_AMD_STRIDED_PHASES", None)
        else:
            os.environ["OWL_AMD_STRIDED_PHASES"] = old
    return dx.double().cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_strided_bwd_input_by_phases(case):
    xs, ks, stride, pad, dt = CASES[case]
    print(f"strided backward input case {case} x{xs} w{ks} s{stride} {dt}", flush=True)
    td = TD[dt]
    rng = np.random.default_rng(900 + case)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) / np.sqrt(np.prod(ks[:-1]))
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
    assert aux == "phase_interleave_kernel", aux
    assert dx.shape == ref.shape
    assert rel(dx, ref) < TOL[dt], rel(dx, ref)
    dg = _bwd_input(xt, wt, stride, gt, generic=True)
    assert rel(dx, dg) < TOL[dt], rel(dx, dg)
    # the same call twice gives the same bits
    assert np.array_equal(dx, _bwd_input(xt, wt, stride, gt, generic=False))


# Forward of strided convolutions whose dilation is a multiple of the stride:
# they read one stride phase of x, gathered, then a stride-1 convolution with
# rate r / s.  (x shape, kernel shape, strides, rates, padding, dtype)
FWD_CASES = [
    ([4, 28, 28, 32], [3, 3, 32, 48], [2, 2], [2, 2], SAME, "bf16"),
    ([2, 29, 31, 16], [5, 5, 16, 24], [2, 2], [2, 2], VALID, "f32"),
    ([2, 16, 17, 8], [3, 3, 8, 16], [3, 3], [3, 3], SAME, "f16"),
    ([2, 9, 10, 16], [3, 3, 16, 8], [2, 1], [2, 1], SAME, "f64"),     # stride 1 along w
    ([2, 14, 12, 16], [3, 3, 16, 16], [2, 2], [4, 2], SAME, "bf16"),  # rate twice the stride
    ([2, 8, 9, 10, 16], [3, 3, 3, 16, 8], [2, 2, 2], [2, 2, 2], SAME, "bf16"),  # 3-D
]


def _fwd(x, w, stride, rate, pad, generic):
    old = os.environ.get("OWL_AMD_STRIDED_FWD_PHASE")
    if generic:
        os.environ["OWL_AMD_STRIDED_FWD_PHASE"] = "0"
    try:
        f = nd.dilated_conv2d if x.dim() == 4 else nd.dilated_conv3d
        y = f(x, w, stride, rate, pad)
        torch.cuda.synchronize()
    finally:
        if old is None:
            os.environ.pop("OWL_AMD_STRIDED_FWD_PHASE", None)
        else:
            os.environ["OWL_AMD_STRIDED_FWD_PHASE"] = old
    return y.double().cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(FWD_CASES)))
def test_strided_dilated_fwd_on_one_phase(case):
    xs, ks, stride, rate, pad, dt = FWD_CASES[case]
    print(f"one-phase forward case {case} x{xs} w{ks} s{stride} r{rate} {dt}", flush=True)
    td = TD[dt]
    rng = np.random.default_rng(950 + case)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) / np.sqrt(np.prod(ks[:-1]))
    rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
    ref = on.dilated_conv2d(rnd(x), rnd(w), stride, rate, pad)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(td)
    xt, wt = dev(x), dev(w)
    _lib.kernel_timing(True)
    y = _fwd(xt, wt, stride, rate, pad, generic=False)
    aux = _lib.kernel_time(_lib.KT_AUX)[2]
    _lib.kernel_timing(False)
    assert aux == "stride_gather_kernel", aux
    assert y.shape == ref.shape
    assert rel(y, ref) < TOL[dt], rel(y, ref)
    yg = _fwd(xt, wt, stride, rate, pad, generic=True)
    assert rel(y, yg) < TOL[dt], rel(y, yg)
    assert np.array_equal(y, _fwd(xt, wt, stride, rate, pad, generic=False))


# Forward of strided 2-D convolutions (unit dilation, 16-bit) as a stride-1
# convolution of x's stride phases stacked as channels (space-to-depth), for
# kernels whose zero-padded phase taps cost at most 1.5x the MACs.
S2D_CASES = [
    ([4, 28, 28, 32], [5, 5, 32, 48], [2, 2], SAME, "bf16"),
    ([2, 29, 31, 16], [5, 5, 16, 24], [2, 2], VALID, "f16"),
    ([2, 30, 30, 8], [7, 7, 8, 16], [2, 2], SAME, "bf16"),
    ([2, 31, 34, 16], [5, 5, 16, 32], [3, 3], SAME, "bf16"),   # stride 3 (pads 2: 2 x 2 phase taps)
    ([2, 20, 22, 16], [5, 3, 16, 16], [2, 1], SAME, "bf16"),   # strided along h only
]


def _fwd_s2d(x, w, stride, pad, generic, bias=None):
    # these shapes are below the sizes where the library takes the path by
    # itself: OWL_AMD_S2D=2 takes it wherever it applies
    old = os.environ.get("OWL_AMD_S2D")
    os.environ["OWL_AMD_S2D"] = "0" if generic else "2"
    try:
        if bias is None:
            y = nd.conv2d(x, w, stride, pad)
        else:
            y = nd.conv2d_bias_act(x, w, bias, stride, pad)
        torch.cuda.synchronize()
    finally:
        if old is None:
            os.environ.pop("OWL_AMD_S2D", None)
        else:
            os.environ["OWL_AMD_S2D"] = old
    return y.double().cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(S2D_CASES)))
def test_strided_fwd_space_to_depth(case):
    xs, ks, stride, pad, dt = S2D_CASES[case]
    print(f"space-to-depth forward case {case} x{xs} w{ks} s{stride} {dt}", flush=True)
    td = TD[dt]
    rng = np.random.default_rng(980 + case)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) / np.sqrt(np.prod(ks[:-1]))
    rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
    ref = on.conv2d(rnd(x), rnd(w), stride, pad)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(td)
    xt, wt = dev(x), dev(w)
    _lib.kernel_timing(True)
    y = _fwd_s2d(xt, wt, stride, pad, generic=False)
    aux = _lib.kernel_time(_lib.KT_AUX)[2]
    fwd = _lib.kernel_time(_lib.KT_FWD)[2]
    _lib.kernel_timing(False)
    assert aux == "space_to_depth_kernel" and fwd == "conv_halo_kernel", (aux, fwd)
    assert y.shape == ref.shape
    assert rel(y, ref) < TOL[dt], rel(y, ref)
    yg = _fwd_s2d(xt, wt, stride, pad, generic=True)
    assert rel(y, yg) < TOL[dt], rel(y, yg)
    assert np.array_equal(y, _fwd_s2d(xt, wt, stride, pad, generic=False))
    # bias and relu pass through to the stride-1 convolution's epilogue
    b = torch.from_numpy(rng.standard_normal(ks[-1]) * 0.1).float().to(DEV)
    yb = _fwd_s2d(xt, wt, stride, pad, generic=False, bias=b)
    ref_b = np.maximum(ref + b.double().cpu().numpy(), 0.0)
    assert rel(yb, ref_b) < TOL[dt], rel(yb, ref_b)


# Forwards with few output tiles over a long K run as split-K on the implicit
# GEMM (FP32 / FP64 partials, the deterministic split-K reduction).
SPLITK_CASES = [
    ([16, 14, 14, 512], [5, 5, 512, 32], [1, 1], [2, 2], VALID, "bf16"),   # 576 outputs, K 12800
    ([5, 7, 7, 512], [3, 3, 512, 256], [2, 2], [2, 2], VALID, "bf16"),     # through the one-phase forward
    ([2, 9, 9, 256], [3, 3, 256, 64], [1, 1], [1, 1], SAME, "f32"),
    ([3, 6, 6, 512], [3, 3, 512, 48], [1, 1], [1, 1], SAME, "f64"),
]


def _fwd_split(x, w, stride, rate, pad, generic):
    old = os.environ.get("OWL_AMD_FWD_SPLITK")
    if generic:
        os.environ["OWL_AMD_FWD_SPLITK"] = "0"
    try:
        y = nd.dilated_conv2d(x, w, stride, rate, pad)
        torch.cuda.synchronize()
    finally:
        if old is None:
            os.environ.pop("OWL_AMD_FWD_SPLITK", None)
        else:
            os.environ["OWL_AMD_FWD_SPLITK"] = old
    return y.double().cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(SPLITK_CASES)))
def test_forward_split_k_for_few_tiles(case):
    xs, ks, stride, rate, pad, dt = SPLITK_CASES[case]
    print(f"split-K forward case {case} x{xs} w{ks} s{stride} r{rate} {dt}", flush=True)
    td = TD[dt]
    rng = np.random.default_rng(990 + case)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) / np.sqrt(np.prod(ks[:-1]))
    rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
    ref = on.dilated_conv2d(rnd(x), rnd(w), stride, rate, pad)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(td)
    xt, wt = dev(x), dev(w)
    _lib.kernel_timing(True)
    y = _fwd_split(xt, wt, stride, rate, pad, generic=False)
    fwd = _lib.kernel_time(_lib.KT_FWD)[2]
    aux = _lib.kernel_time(_lib.KT_AUX)[2]
    _lib.kernel_timing(False)
    assert fwd == "conv_igemm_kernel" and aux == "splitk_reduce_kernel", (fwd, aux)
    assert y.shape == ref.shape
    assert rel(y, ref) < TOL[dt], rel(y, ref)
    yg = _fwd_split(xt, wt, stride, rate, pad, generic=True)
    assert rel(y, yg) < TOL[dt], rel(y, yg)
    assert np.array_equal(y, _fwd_split(xt, wt, stride, rate, pad, generic=False))
