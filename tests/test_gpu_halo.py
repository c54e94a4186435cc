"""Stride-1 2-D problems that take the LDS-halo kernel (16-bit storage,
channels a multiple of 32, output channels a multiple of 8, rows up to 256
pixels), checked against the FP64 oracle on the same rounded inputs.

The shapes cover both halo modes: halos of at most 352 pixels are double
buffered (the next channel block loads while the current one runs); larger
ones (112-pixel rows, 5x5 kernels) restage a single buffer between channel
blocks.  Also 1x1 kernels (one tap per block), dilation, VALID padding and
output channels that do not fill a tile (the forward reads the weights in
place with transposing LDS loads; columns past the last channel read as
zeros).  FP32 output: rel-L2 < 1e-5 (the MFMA accumulation is the only
error); BF16 output: rel-L2 < 1e-2 end to end.

Each case runs with the fragment edge the library picks and with each edge
forced (OWL_AMD_HALO_MS): 32-edge tiles (256 / 512 pixels, 32x32x16 MFMAs)
and 16-edge tiles (224 / 448 pixels, 16x16x32 MFMAs, 2 x 2 waves for
128-wide channel slices), including short last row tiles, a partial second
channel slice and the backward input with 128 input channels."""
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
