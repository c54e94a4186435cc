"""Degenerate shapes whose results are all zeros (Owl's C memsets its outputs
first: owl_ndarray_conv_impl.h:285, 719; owl_ndarray_upsampling_impl.h:20):
a weight gradient over an empty batch, the bias gradient of an empty layer.
The library clears them with a fill kernel, so the results are zeros at any
byte alignment and a HIP graph that captures such a call clears them again on
every replay."""
import ctypes

import pytest

from owl_amd import _lib
from owl_amd import ndarray as nd

torch = pytest.importorskip("torch")
DEV = "cuda:0"


def _bwd_kernel_empty_batch(dw_view):
    # x [0, 5, 6, 3] (*) w [2, 3, 3, 4]: K = B * Ho * Wo = 0
    g = nd._Conv(0, [5, 6], 3, [2, 3], 4, [5, 6], [1, 1], [1, 1], [0, 1])
    d = g.desc()
    x = torch.empty(16, device=DEV)
    dy = torch.empty(16, device=DEV)
    _lib.check(_lib.lib().owl_amd_conv_bwd_kernel(
        ctypes.byref(d), _lib.F32, _lib.F32, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(dy.data_ptr()),
        ctypes.c_void_p(dw_view.data_ptr()), None, ctypes.c_size_t(0),
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))


@pytest.mark.gpu
@pytest.mark.parametrize("offset", [0, 1, 3, 5])
def test_empty_batch_weight_gradient_is_zero(offset):
    n = 2 * 3 * 3 * 4
    buf = torch.full((n + 16,), float("nan"), device=DEV)
    view = buf[offset:offset + n]          # 4-byte steps: unaligned heads and tails
    _bwd_kernel_empty_batch(view)
    torch.cuda.synchronize()
    assert torch.equal(view, torch.zeros_like(view))
    assert torch.isnan(buf[:offset]).all() and torch.isnan(buf[offset + n:]).all()


@pytest.mark.gpu
def test_empty_layer_bias_gradient_is_zero():
    dy = torch.empty(0, 7, device=DEV, dtype=torch.bfloat16)
    pre = torch.empty_like(dy)
    dz, db = nd.bias_act_backward(pre, dy)
    torch.cuda.synchronize()
    assert dz.numel() == 0 and torch.equal(db, torch.zeros(7, device=DEV))


@pytest.mark.gpu
def test_captured_zero_fill_clears_on_every_replay():
    n = 2 * 3 * 3 * 4
    buf = torch.full((n + 8,), float("nan"), device=DEV)
    view = buf[3:3 + n]
    _bwd_kernel_empty_batch(view)           # outside the capture first
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        _bwd_kernel_empty_batch(view)
    torch.cuda.synchronize()
    for rep in range(3):
        view.fill_(float(rep + 1))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(view, torch.zeros_like(view)), rep
