"""The Owl drop-ins on host (numpy) buffers: arrays larger than one pinned
staging chunk (8 MiB) move through several chunks and give the same bits as
the device path on the same data; calls from several host threads at once
(ctypes releases the GIL) each get their own staging context and give the
same bits as the same calls made one after another, with the drop-in graph
cache off and on (strided backward inputs among them, which share
per-device buffers and so run under the library lock); repeated calls of one shape replay captured graphs with
the bits of the device path."""
import threading

import numpy as np
import pytest

from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")
DEV = "cuda:0"


@pytest.mark.gpu
def test_multi_chunk_host_arrays_equal_device_path():
    rng = np.random.default_rng(3)
    x = rng.standard_normal([12, 56, 56, 64]).astype(np.float32)          # 9.6 MB: two chunks
    w = (rng.standard_normal([3, 3, 64, 128]) * 0.06).astype(np.float32)
    y = nd.conv2d(x, w, [1, 1], SAME)                                        # 19.3 MB: three chunks
    dy = rng.standard_normal(y.shape).astype(np.float32)
    dx = nd.conv2d_backward_input(x, w, [1, 1], dy)
    dw = nd.conv2d_backward_kernel(x, w, [1, 1], dy)
    xt, wt, dyt = (torch.from_numpy(a).to(DEV) for a in (x, w, dy))
    assert np.array_equal(y, nd.conv2d(xt, wt, [1, 1], SAME).cpu().numpy())
    assert np.array_equal(dx, nd.conv2d_backward_input(xt, wt, [1, 1], dyt).cpu().numpy())
    assert np.array_equal(dw, nd.conv2d_backward_kernel(xt, wt, [1, 1], dyt).cpu().numpy())
    p, a = nd.max_pool2d_argmax(y, [2, 2], [2, 2], VALID)
    pt, at = nd.max_pool2d_argmax(torch.from_numpy(y).to(DEV), [2, 2], [2, 2], VALID)
    assert np.array_equal(p, pt.cpu().numpy()) and np.array_equal(a, at.cpu().numpy())


def _work(seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal([2 + seed % 3, 20, 18, 16]).astype(np.float32)
    w = (rng.standard_normal([3, 3, 16, 24]) * 0.1).astype(np.float32)
    y = nd.conv2d(x, w, [1, 1], SAME)
    dy = rng.standard_normal(y.shape).astype(np.float32)
    out = [y, nd.conv2d_backward_input(x, w, [1, 1], dy), nd.conv2d_backward_kernel(x, w, [1, 1], dy)]
    p = nd.max_pool2d(y, [3, 3], [2, 2], SAME)
    out.append(nd.max_pool2d_backward(SAME, y, [3, 3], [2, 2], rng.standard_normal(p.shape).astype(np.float32)))
    # a strided 1x1 backward input: its subsampled grid goes through one
    # per-device buffer, whose size differs between the threads' calls
    w1 = (rng.standard_normal([1, 1, 16, 32 + 8 * (seed % 3)]) * 0.2).astype(np.float32)
    y1 = nd.conv2d(x, w1, [2, 2], SAME)
    out.append(nd.conv2d_backward_input(x, w1, [2, 2], rng.standard_normal(y1.shape).astype(np.float32)))
    # and a strided 3x3 one (stride phases: phase grids and grouped taps in
    # per-device buffers)
    y3 = nd.conv2d(x, w, [2, 2], SAME)
    out.append(nd.conv2d_backward_input(x, w, [2, 2], rng.standard_normal(y3.shape).astype(np.float32)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [0, 1])
def test_concurrent_dropin_calls_equal_sequential(graphs):
    from owl_amd import _lib
    prev = _lib.lib().owl_amd_dropin_graphs(graphs)
    try:
        _check_concurrent()
    finally:
        _lib.lib().owl_amd_dropin_graphs(prev)


def _check_concurrent():
    seeds = list(range(6))
    ref = {s: _work(s) for s in seeds}
    got, errors = {}, []

    def run(s):
        try:
            for _ in range(3):
                got[s] = _work(s)
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)
    threads = [threading.Thread(target=run, args=(s,)) for s in seeds]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for s in seeds:
        for a, b in zip(got[s], ref[s]):
            assert np.array_equal(a, b), s


def _graph_stats():
    import ctypes
    from owl_amd import _lib
    g, c, r = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.lib().owl_amd_dropin_graph_stats(ctypes.byref(g), ctypes.byref(c), ctypes.byref(r)))
    return g.value, c.value, r.value


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_repeated_dropin_calls_replay_graphs(dtype):
    """With the drop-in graph cache on, four calls of each drop-in on one shape with fresh data each time (the
    first runs directly, the second is captured, the rest replay the graph):
    every result equals the device path on the same data, bit for bit --
    including a split-K weight gradient, max pooling with an all-NaN window,
    the argmax and the upsampling."""
    from owl_amd import _lib
    prev = _lib.lib().owl_amd_dropin_graphs(1)
    try:
        _check_repeated_calls(dtype)
    finally:
        _lib.lib().owl_amd_dropin_graphs(prev)


def _check_repeated_calls(dtype):
    _, c0, r0 = _graph_stats()
    for call in range(4):
        rng = np.random.default_rng(100 + call)
        x = rng.standard_normal([4, 30, 30, 32]).astype(dtype)
        w = (rng.standard_normal([3, 3, 32, 48]) * 0.08).astype(dtype)
        y0 = nd.conv2d(x, w, [1, 1], SAME)
        dy = rng.standard_normal(y0.shape).astype(dtype)
        dx = nd.conv2d_backward_input(x, w, [1, 1], dy)
        dw = nd.conv2d_backward_kernel(x, w, [1, 1], dy)
        y = y0.copy()
        y[1, 4:6, 6:8, 3] = np.nan          # one 2x2 window of NaNs
        p = nd.max_pool2d(y, [2, 2], [2, 2], VALID)
        dp = rng.standard_normal(p.shape).astype(dtype)
        dmx = nd.max_pool2d_backward(VALID, y, [2, 2], [2, 2], dp)
        dav = nd.avg_pool2d_backward(VALID, y, [2, 2], [2, 2], dp)
        pa, arg = nd.max_pool2d_argmax(y, [2, 2], [2, 2], VALID)
        up = nd.upsampling2d(p, [2, 2])
        xt, wt, dyt, yt, dpt = (torch.from_numpy(a).to(DEV) for a in (x, w, dy, y, dp))
        assert np.array_equal(y0, nd.conv2d(xt, wt, [1, 1], SAME).cpu().numpy())
        assert np.array_equal(dx, nd.conv2d_backward_input(xt, wt, [1, 1], dyt).cpu().numpy())
        assert np.array_equal(dw, nd.conv2d_backward_kernel(xt, wt, [1, 1], dyt).cpu().numpy())
        assert np.array_equal(p, nd.max_pool2d(yt, [2, 2], [2, 2], VALID).cpu().numpy(), equal_nan=True)
        assert np.array_equal(dmx, nd.max_pool2d_backward(VALID, yt, [2, 2], [2, 2], dpt).cpu().numpy())
        assert np.array_equal(dav, nd.avg_pool2d_backward(VALID, yt, [2, 2], [2, 2], dpt).cpu().numpy())
        pat, argt = nd.max_pool2d_argmax(yt, [2, 2], [2, 2], VALID)
        assert np.array_equal(pa, pat.cpu().numpy(), equal_nan=True) and np.array_equal(arg, argt.cpu().numpy())
        assert np.array_equal(up, nd.upsampling2d(torch.from_numpy(p).to(DEV), [2, 2]).cpu().numpy(),
                              equal_nan=True)
    _, c1, r1 = _graph_stats()
    # eight call kinds on host arrays, each captured on its second call and
    # replayed on the later ones (a context per call kind at most: more
    # captures only when calls land on other staging contexts)
    assert c1 - c0 >= 8 and r1 - r0 >= 2 * 8, (c0, c1, r0, r1)
