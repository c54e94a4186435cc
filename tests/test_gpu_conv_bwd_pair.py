"""conv2d_backward / conv3d_backward: both backward passes in one call (the
pair Owl's reverse rule computes).  The results have the bits of the two
separate calls; torch work queued right after the call reads the final dw;
and a HIP graph that captures the call replays it with the same bits, with
eager work between the replays."""
import pytest

from owl_amd import ndarray as nd

torch = pytest.importorskip("torch")
DEV = "cuda:0"


def _inputs(xs, ks, stride, dt, seed):
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    x = torch.randn(xs, device=DEV, generator=g).to(dt)
    w = (torch.randn(ks, device=DEV, generator=g) * 0.05).to(dt)
    geo = nd.conv_problem(xs, ks, stride, None, "SAME")
    dy = torch.randn([xs[0], *geo.sp_out, ks[-1]], device=DEV, generator=g).to(dt)
    return x, w, dy


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [torch.bfloat16, torch.float32, torch.float64], ids=["bf16", "fp32", "fp64"])
@pytest.mark.parametrize("case", [([16, 28, 28, 64], [3, 3, 64, 128], [1, 1]),
                                  ([4, 17, 13, 24], [3, 2, 24, 40], [2, 1]),
                                  ([2, 9, 10, 11, 16], [3, 3, 3, 16, 24], [1, 1, 1])], ids=["vgg", "odd", "3d"])
def test_pair_equals_separate_calls(dt, case):
    xs, ks, stride = case
    x, w, dy = _inputs(xs, ks, stride, dt, 5)
    wdt = torch.float64 if dt == torch.float64 else torch.float32
    bi, bk, both = ((nd.conv2d_backward_input, nd.conv2d_backward_kernel, nd.conv2d_backward) if len(xs) == 4 else
                    (nd.conv3d_backward_input, nd.conv3d_backward_kernel, nd.conv3d_backward))
    rdx = bi(x, w, stride, dy)
    rdw = bk(x, w, stride, dy, out_dtype=wdt)
    dx, dw = both(x, w, stride, dy, kernel_out_dtype=wdt)
    s = dw.sum()           # queued right after the call on the same stream
    torch.cuda.synchronize()
    assert torch.equal(dx, rdx) and torch.equal(dw, rdw)
    assert torch.equal(s, rdw.sum())


@pytest.mark.gpu
def test_pair_under_graph_capture():
    x, w, dy = _inputs([8, 28, 28, 64], [3, 3, 64, 128], [1, 1], torch.bfloat16, 6)
    rdx = nd.conv2d_backward_input(x, w, [1, 1], dy)
    rdw = nd.conv2d_backward_kernel(x, w, [1, 1], dy, out_dtype=torch.float32)
    need = nd._ws_bytes(nd._bwd_geom(2, list(x.shape), list(w.shape), [1, 1], None, list(dy.shape)), x)
    ws = torch.empty(max(need, 1), dtype=torch.uint8, device=DEV)
    nd.conv2d_backward(x, w, [1, 1], dy, kernel_out_dtype=torch.float32, workspace=ws)  # warm
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        dx, dw = nd.conv2d_backward(x, w, [1, 1], dy, kernel_out_dtype=torch.float32, workspace=ws)
    torch.cuda.synchronize()
    for rep in range(3):
        dx.zero_()
        dw.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(dx, rdx) and torch.equal(dw, rdw), rep
        # eager paired calls between the replays
        x2, w2, dy2 = _inputs([2, 14, 14, 64], [3, 3, 64, 128], [1, 1], torch.bfloat16, 10 + rep)
        nd.conv2d_backward(x2, w2, [1, 1], dy2)
        torch.cuda.synchronize()
