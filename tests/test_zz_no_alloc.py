"""owl_amd_conv_bwd_kernel_noalloc: the backward-kernel pass with its split-K
partials in the library's file-scope device scratch instead of an allocated
workspace.  Through the C ABI: two calls in a row (the second smaller) give
the same bits as owl_amd_conv_bwd_kernel with a caller workspace on the same
inputs (same kernels, same split plan), in BF16 and FP32; the call creates no
library buffer; every split plan's partials fit the scratch; and
a HIP graph that captures the call replays it with the same bits."""
import ctypes

import pytest

from owl_amd import _lib
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME

torch = pytest.importorskip("torch")
DEV = "cuda:0"
LIMIT = 32 * 1024 * 1024


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _problem(xs, ks, stride, dt, seed):
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    x = torch.randn(xs, device=DEV, generator=g).to(dt)
    geo = nd.conv_problem(xs, ks, stride, None, SAME)
    dy = torch.randn([xs[0], *geo.sp_out, ks[3]], device=DEV, generator=g).to(dt)
    return x, dy, geo


def _noalloc(geo, x, dy, dw):
    d = geo.desc()
    return _lib.lib().owl_amd_conv_bwd_kernel_noalloc(
        ctypes.byref(d), nd._TORCH_DT[x.dtype], nd._TORCH_DT[dw.dtype], ctypes.c_void_p(x.data_ptr()),
        ctypes.c_void_p(dy.data_ptr()), ctypes.c_void_p(dw.data_ptr()), _stream())


def _with_workspace(geo, x, dy, dw, ws, need):
    d = geo.desc()
    return _lib.lib().owl_amd_conv_bwd_kernel(
        ctypes.byref(d), nd._TORCH_DT[x.dtype], nd._TORCH_DT[dw.dtype], ctypes.c_void_p(x.data_ptr()),
        ctypes.c_void_p(dy.data_ptr()), ctypes.c_void_p(dw.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
        ctypes.c_size_t(need), _stream())


def _need(geo, dt):
    d = geo.desc()
    return _lib.lib().owl_amd_conv_bwd_kernel_workspace(ctypes.byref(d), nd._TORCH_DT[dt])


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [torch.bfloat16, torch.float32], ids=["bf16", "fp32"])
def test_noalloc_equals_workspace_entry_twice(dt):
    big = [64, 56, 56, 64] if dt == torch.bfloat16 else [16, 56, 56, 64]
    cases = [(big, [3, 3, 64, 128], [1, 1]),           # a benchmark-sized layer
             ([6, 40, 44, 40], [3, 2, 40, 24], [2, 1])]  # then a smaller one, stride 2
    _lib.lib().owl_amd_cache_release()
    for i, (xs, ks, stride) in enumerate(cases):
        x, dy, geo = _problem(xs, ks, stride, dt, 70 + i)
        need = _need(geo, dt)
        assert 0 < need <= LIMIT, need
        ws = torch.empty(need, dtype=torch.uint8, device=DEV)
        ref = torch.empty(ks, dtype=torch.float32, device=DEV)
        _lib.check(_with_workspace(geo, x, dy, ref, ws, need))
        dw = torch.full(ks, float("nan"), dtype=torch.float32, device=DEV)
        _lib.check(_noalloc(geo, x, dy, dw))
        torch.cuda.synchronize()
        assert torch.equal(dw, ref), (i, (dw - ref).abs().max().item())
        n, b = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib().owl_amd_cache_stats(ctypes.byref(n), ctypes.byref(b)))
        assert b.value == 0, b.value      # no library buffer was created


def test_split_plans_fit_the_scratch():
    """The split plan bounds the partials by one round of workgroups' tiles
    (at most 256 FP32 tiles of 192 x 128 for the weight-gradient kernel, 512
    FP64 tiles of 64 x 64 for the implicit GEMM), so every real problem's
    partials fit the scratch and the entry point's size check (an ERR_ARG
    naming the scratch) is a guard no plan reaches.  Host-side query only."""
    import itertools
    worst = {}
    for dt in (torch.float32, torch.float64, torch.bfloat16, torch.float16):
        for b, h, c, co, k, s in itertools.product([1, 8, 64, 256], [7, 28, 56, 112], [16, 64, 256, 512],
                                                   [16, 128, 512], [1, 3, 5], [1, 2]):
            geo = nd.conv_problem([b, h, h, c], [k, k, c, co], [s, s], None, SAME)
            worst[dt] = max(worst.get(dt, 0), _need(geo, dt))
    assert all(0 < v <= LIMIT for v in worst.values()), worst


@pytest.mark.gpu
def test_noalloc_under_graph_capture():
    x, dy, geo = _problem([8, 28, 28, 64], [3, 3, 64, 64], [1, 1], torch.bfloat16, 90)
    w = torch.empty([3, 3, 64, 64], dtype=torch.bfloat16, device=DEV)
    ref = nd.conv2d_backward_kernel(x, w, [1, 1], dy, out_dtype=torch.float32)
    dw = torch.empty_like(ref)
    _lib.check(_noalloc(geo, x, dy, dw))
    torch.cuda.synchronize()
    assert torch.equal(dw, ref)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        _lib.check(_noalloc(geo, x, dy, dw))
    torch.cuda.synchronize()
    for rep in range(3):
        dw.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(dw, ref), rep
        # eager calls on the same scratch between the replays
        x2, dy2, g2 = _problem([2, 14, 14, 64], [3, 3, 64, 64], [1, 1], torch.bfloat16, 100 + rep)
        o2 = torch.empty([3, 3, 64, 64], dtype=torch.float32, device=DEV)
        _lib.check(_noalloc(g2, x2, dy2, o2))
        torch.cuda.synchronize()
