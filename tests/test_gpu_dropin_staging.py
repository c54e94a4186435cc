"""The Owl drop-ins on host (numpy) buffers: arrays larger than one pinned
staging chunk (8 MiB) move through several chunks and give the same bits as
the device path on the same data; calls from several host threads at once
(ctypes releases the GIL) each get their own staging context and give the
same bits as the same calls made one after another."""
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
    return out


@pytest.mark.gpu
def test_concurrent_dropin_calls_equal_sequential():
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
