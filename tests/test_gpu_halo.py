"""Stride-1 2-D and 3-D problems that take the LDS-halo kernel (16-bit
storage, channels a multiple of 8 -- those not a multiple of 32 on the
general instance, whose last channel block reads zeros past the last
channel --, output channels a multiple of 8, rows up to 256 pixels),
checked against the FP64 oracle on the same rounded inputs.

The 2-D shapes cover halos up to 456 pixels (112-pixel rows, 5x5 kernels),
restaged between channel blocks.  Also 1x1 kernels (one tap per block), dilation, VALID padding and
output channels that do not fill a tile (the forward reads the weights in
place with transposing LDS loads; columns past the last channel read as
zeros).  FP32 output: rel-L2 < 1e-5 (the MFMA accumulation is the only
error); BF16 output: rel-L2 < 1e-2 end to end.

Each case runs with the fragment edge the library picks and with each edge
forced (OWL_AMD_HALO_MS): 32-edge tiles (256 / 512 pixels, 32x32x16 MFMAs)
and 16-edge tiles (224 / 448 pixels, 16x16x32 MFMAs, 2 x 2 waves for
128-wide channel slices), including short last row tiles, a partial second
channel slice and the backward input with 128 input channels.  With the
edge left to the library, the backward input of layers with more than 64
input channels takes 64-wide slices where their 448-pixel tiles fill the
rows as well as 224-pixel ones do (the last two 2-D cases before the
3-D ones, and the last 3-D case); a forced edge keeps the default slice.

3-D layers run the same kernel over (w, d) planes, with the kernel's h taps
as outer blocks that restage the halo from another input plane (zeros past
the image): SAME and VALID, dilation, even and one-wide kernel dims, a
partial second channel slice; the library's kernel timing confirms that the
halo kernel ran."""
import os

import numpy as np
import pytest

import oracle_ndarray as on
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")

DEV = "cuda:0"

# (x shape, kernel shape, rate, padding)
CASES = [
    ([2, 12, 56, 64], [3, 3, 64, 128], 1, SAME),     # double buffered, 2 channel blocks
    ([1, 6, 112, 32], [3, 3, 32, 64], 1, SAME),      # 456-pixel halo: single buffer
    ([1, 5, 112, 96], [3, 3, 96, 40], 1, SAME),      # single buffer, 3 blocks, partial N tile
    ([2, 9, 28, 64], [5, 5, 64, 32], 1, SAME),       # 5x5: 416-pixel halo
    ([2, 10, 20, 32], [3, 3, 32, 64], 2, VALID),     # dilated
    ([1, 7, 7, 256], [1, 1, 256, 72], 1, VALID),     # one tap, 8 blocks
    ([3, 14, 14, 128], [3, 3, 128, 128], 1, SAME),   # whole image per tile
    ([2, 8, 30, 32], [3, 3, 32, 200], 1, SAME),      # 128-wide slices, the last one partial
    ([3, 13, 56, 64], [3, 3, 64, 128], 1, SAME),     # row tiles 4,4,4,1 (forward), 8,5 (backward input)
    ([2, 11, 30, 64], [3, 3, 64, 200], 1, SAME),     # two 128-wide slices, the second 72 wide
    ([4, 14, 14, 128], [3, 3, 128, 256], 1, SAME),   # backward input with 128 input channels
    ([2, 10, 30, 16], [3, 3, 16, 32], 1, SAME),      # 16 channels: one zero-padded block (general instance)
    ([1, 8, 20, 40], [3, 3, 40, 24], 1, SAME),       # 40 / 24 channels: a partial last block
    ([2, 16, 56, 128], [3, 3, 128, 64], 1, SAME),    # backward input (N = 128) on two 64-wide slices
    ([1, 32, 28, 96], [3, 3, 96, 32], 1, SAME),      # backward input (N = 96) on 64-wide slices, the second half used
]


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _run(x, w, g, rate, pad, td, out_dtype):
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(td)
    xt, wt, gt = dev(x), dev(w), dev(g)
    s, r = [1, 1], [rate, rate]
    y = nd.dilated_conv2d(xt, wt, s, r, pad, out_dtype=out_dtype)
    dx = nd.dilated_conv2d_backward_input(xt, wt, s, r, gt, out_dtype=out_dtype)
    torch.cuda.synchronize()
    return y.double().cpu().numpy(), dx.double().cpu().numpy()


@pytest.fixture(params=["auto", "16", "32"])
def halo_edge(request):
    old = os.environ.get("OWL_AMD_HALO_MS")
    if request.param == "auto":
        os.environ.pop("OWL_AMD_HALO_MS", None)
    else:
        os.environ["OWL_AMD_HALO_MS"] = request.param
    yield request.param
    if old is None:
        os.environ.pop("OWL_AMD_HALO_MS", None)
    else:
        os.environ["OWL_AMD_HALO_MS"] = old


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_halo_vs_oracle(case, dtype, halo_edge):
    xs, ks, rate, pad = CASES[case]
    print(f"halo case {case} {dtype} edge {halo_edge} x{xs} w{ks}", flush=True)
    td = torch.bfloat16 if dtype == "bf16" else torch.float16
    rng = np.random.default_rng(100 + case)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) / np.sqrt(ks[0] * ks[1] * ks[2])
    ys = on.dilated_conv2d(x, w, [1, 1], [rate, rate], pad).shape
    g = rng.standard_normal(ys)
    rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
    xr, wr, gr = rnd(x), rnd(w), rnd(g)
    ry = on.dilated_conv2d(xr, wr, [1, 1], [rate, rate], pad)
    rdx = on.dilated_conv2d_backward_input(xr, wr, [1, 1], [rate, rate], gr)
    y, dx = _run(x, w, g, rate, pad, td, torch.float32)
    assert rel(y, ry) < 1e-5, rel(y, ry)
    assert rel(dx, rdx) < 1e-5, rel(dx, rdx)
    y, dx = _run(x, w, g, rate, pad, td, None)
    assert rel(y, ry) < 1e-2 and rel(dx, rdx) < 1e-2, (rel(y, ry), rel(dx, rdx))
    # the same call twice gives the same bits (no state left between calls)
    y2, dx2 = _run(x, w, g, rate, pad, td, None)
    assert np.array_equal(y, y2) and np.array_equal(dx, dx2)


# 3-D: (x shape, kernel shape, rate, padding); both channel counts multiples of 8,
# so the forward and the backward input both take the halo kernel
CASES3 = [
    ([2, 6, 20, 18, 32], [3, 3, 3, 32, 64], 1, SAME),     # the conv3d stack's second layer, smaller
    ([1, 5, 12, 14, 64], [3, 3, 3, 64, 32], 1, SAME),     # two channel blocks per outer tap
    ([2, 7, 9, 30, 32], [3, 3, 3, 32, 64], 2, VALID),     # dilated, VALID
    ([1, 4, 8, 40, 32], [2, 3, 5, 32, 32], 1, SAME),      # even kh, uneven kw / kd
    ([2, 8, 16, 16, 32], [3, 3, 3, 32, 160], 1, SAME),    # two 128-wide slices, the second partial
    ([1, 3, 5, 6, 32], [3, 1, 1, 32, 32], 1, SAME),       # only outer taps
    ([2, 6, 12, 10, 16], [3, 3, 3, 16, 32], 1, SAME),     # 16 input channels (the conv3d stack's first layer)
    ([1, 5, 8, 9, 24], [3, 3, 3, 24, 40], 1, SAME),       # 24 / 40 channels: partial blocks
    ([1, 4, 16, 28, 96], [3, 3, 3, 96, 32], 1, SAME),     # backward input (N = 96) on 64-wide slices
]


def _run3(x, w, g, rate, pad, td, out_dtype):
    from owl_amd import _lib
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(td)
    xt, wt, gt = dev(x), dev(w), dev(g)
    s, r = [1, 1, 1], [rate, rate, rate]
    torch.cuda.synchronize()
    _lib.kernel_timing(True)
    y = nd.dilated_conv3d(xt, wt, s, r, pad, out_dtype=out_dtype)
    dx = nd.dilated_conv3d_backward_input(xt, wt, s, r, gt, out_dtype=out_dtype)
    torch.cuda.synchronize()
    names = (_lib.kernel_time(_lib.KT_FWD)[2], _lib.kernel_time(_lib.KT_BWD_INPUT)[2])
    _lib.kernel_timing(False)
    return y.double().cpu().numpy(), dx.double().cpu().numpy(), names


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES3)))
@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_halo3d_vs_oracle(case, dtype, halo_edge):
    xs, ks, rate, pad = CASES3[case]
    print(f"halo 3-D case {case} {dtype} edge {halo_edge} x{xs} w{ks}", flush=True)
    td = torch.bfloat16 if dtype == "bf16" else torch.float16
    rng = np.random.default_rng(300 + case)
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) / np.sqrt(ks[0] * ks[1] * ks[2] * ks[3])
    ys = on.dilated_conv2d(x, w, [1, 1, 1], [rate] * 3, pad).shape
    g = rng.standard_normal(ys)
    rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
    xr, wr, gr = rnd(x), rnd(w), rnd(g)
    ry = on.dilated_conv2d(xr, wr, [1, 1, 1], [rate] * 3, pad)
    rdx = on.dilated_conv2d_backward_input(xr, wr, [1, 1, 1], [rate] * 3, gr)
    y, dx, names = _run3(x, w, g, rate, pad, td, torch.float32)
    assert names == ("conv_halo_kernel", "conv_halo_kernel"), names
    assert rel(y, ry) < 1e-5, rel(y, ry)
    assert rel(dx, rdx) < 1e-5, rel(dx, rdx)
    y, dx, _ = _run3(x, w, g, rate, pad, td, None)
    assert rel(y, ry) < 1e-2 and rel(dx, rdx) < 1e-2, (rel(y, ry), rel(dx, rdx))
    y2, dx2, _ = _run3(x, w, g, rate, pad, td, None)
    assert np.array_equal(y, y2) and np.array_equal(dx, dx2)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_small_image_backward_input_leaves_the_halo_kernel(dtype):
    """7x7 images fill 49 of a halo tile's 224 pixels: the 2-D backward input
    runs on the direct-A kernel (tiles of pixels across images) instead, and
    matches the oracle; the forward stays on the halo kernel."""
    from owl_amd import _lib
    td = torch.bfloat16 if dtype == "bf16" else torch.float16
    rng = np.random.default_rng(77)
    xs, ks = [6, 7, 7, 64], [3, 3, 64, 96]
    x = rng.standard_normal(xs)
    w = rng.standard_normal(ks) / np.sqrt(9 * 64)
    g = rng.standard_normal([6, 7, 7, 96])
    rnd = lambda a: torch.from_numpy(a).to(td).to(torch.float64).numpy()
    ry = on.conv2d(rnd(x), rnd(w), [1, 1], SAME)
    rdx = on.conv2d_backward_input(rnd(x), rnd(w), [1, 1], rnd(g))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV).to(td)
    xt, wt, gt = dev(x), dev(w), dev(g)
    _lib.kernel_timing(True)
    y = nd.conv2d(xt, wt, [1, 1], SAME, out_dtype=torch.float32)
    dx = nd.conv2d_backward_input(xt, wt, [1, 1], gt, out_dtype=torch.float32)
    torch.cuda.synchronize()
    names = (_lib.kernel_time(_lib.KT_FWD)[2], _lib.kernel_time(_lib.KT_BWD_INPUT)[2])
    _lib.kernel_timing(False)
    assert names == ("conv_halo_kernel", "conv_fast_kernel"), names
    assert rel(y.double().cpu().numpy(), ry) < 1e-5
    assert rel(dx.double().cpu().numpy(), rdx) < 1e-5
