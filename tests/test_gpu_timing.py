"""The kernel-timing diagnostics: each convolution call records its main
kernel (and the auxiliary ones) with HIP events on the stream it launches on,
only while timing is enabled."""
import pytest

from owl_amd import _lib
from owl_amd import ndarray as nd

torch = pytest.importorskip("torch")


@pytest.mark.gpu
def test_kernel_timing_counts_launches():
    x = torch.randn(2, 16, 16, 64, device="cuda").bfloat16()
    w = (torch.randn(3, 3, 64, 32, device="cuda") * 0.05).bfloat16()
    dy = torch.randn(2, 16, 16, 32, device="cuda").bfloat16()
    _lib.kernel_timing(True)
    for _ in range(3):
        nd.conv2d(x, w, [1, 1])
        nd.conv2d_backward_input(x, w, [1, 1], dy)
        nd.conv2d_backward_kernel(x, w, [1, 1], dy, out_dtype=torch.float32)
    torch.cuda.synchronize()
    got = {c: _lib.kernel_time(c) for c in (_lib.KT_FWD, _lib.KT_BWD_INPUT, _lib.KT_BWD_KERNEL, _lib.KT_AUX)}
    _lib.kernel_timing(False)
    for c in (_lib.KT_FWD, _lib.KT_BWD_INPUT, _lib.KT_BWD_KERNEL):
        ms, n, name = got[c]
        assert n == 3 and ms > 0 and name.startswith("conv_"), (c, got[c])
    # disabled: nothing more is recorded
    nd.conv2d(x, w, [1, 1])
    torch.cuda.synchronize()
    assert _lib.kernel_time(_lib.KT_FWD)[1] == 3
    with pytest.raises(_lib.OwlAmdError):
        _lib.kernel_time(7)
