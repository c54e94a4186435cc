"""16-bit weight gradients that take the halo form of the weight-gradient
kernel (2-D, kw = 3, unit stride along W, input channels a multiple of 64
and at least 128, Wo >= 32, B Ho Wo a multiple of 32): the 32 pixels of a
K-step fall into one or two output rows, across image boundaries,
with SAME / VALID pads, kernel heights 1 to 3 and a stride along H.  Each
case is checked against the FP64 oracle on the same rounded inputs (FP32
output: rel-L2 < 1e-5); the library's kernel timing names the kernel that
ran."""
import numpy as np
import pytest

import oracle_ndarray as on
from owl_amd import _lib
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")
DEV = "cuda:0"

# (x shape, kernel shape, stride, padding); B Ho Wo % 32 == 0 in each
CASES = [
    ([2, 16, 32, 128], [3, 3, 128, 64], [1, 1], SAME),     # Wo 32: every step one whole row
    ([4, 12, 40, 128], [3, 3, 128, 128], [1, 1], SAME),    # Wo 40: runs cross rows and images
    ([2, 20, 48, 192], [2, 3, 192, 40], [1, 1], SAME),     # kh 2, three 64-channel blocks, partial N tile
    ([3, 18, 34, 128], [3, 3, 128, 136], [1, 1], VALID),   # VALID, Wo 32, two 128-wide N tiles
    ([2, 32, 36, 128], [1, 3, 128, 64], [2, 1], SAME),     # kh 1, stride 2 along H
    ([1, 56, 56, 128], [3, 3, 128, 128], [1, 1], SAME),    # the 128-channel VGG layer shape
]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("td", [torch.bfloat16, torch.float16], ids=["bf16", "fp16"])
def test_wgrad_halo_vs_oracle(case, td):
    xs, ks, stride, pad = CASES[case]
    rng = np.random.default_rng(case)
    x = rng.standard_normal(xs).astype(np.float32)
    w = (rng.standard_normal(ks) * 0.1).astype(np.float32)
    geo = nd.conv_problem(xs, ks, stride, None, pad)
    assert (geo.B * geo.sp_out[0] * geo.sp_out[1]) % 32 == 0
    g = rng.standard_normal([xs[0], *geo.sp_out, ks[3]]).astype(np.float32)
    xt, wt, gt = (torch.from_numpy(a).to(DEV).to(td) for a in (x, w, g))
    _lib.kernel_timing(True)
    dw = nd.conv2d_backward_kernel(xt, wt, stride, gt, out_dtype=torch.float32)
    torch.cuda.synchronize()
    _, n, name = _lib.kernel_time(_lib.KT_BWD_KERNEL)
    _lib.kernel_timing(False)
    assert n == 1 and name == "conv_wgrad_halo_kernel", name
    h = lambda t: t.double().cpu().numpy()
    ref = on.conv2d_backward_kernel(h(xt), h(wt), stride, h(gt))
    assert rel(dw.cpu().numpy(), ref) < 1e-5
