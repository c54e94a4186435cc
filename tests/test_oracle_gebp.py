"""Owl's FP32 CPU algorithm as the oracle restates it (the CPU baseline):
the block sizes of compute_block_sizes (owl_ndarray_conv_impl.h:21-87) on
known answers worked out by hand from its formulas, and the GEBP branch
(:340-452, :553-663, :774-887) against the FP64 oracle, including a layer whose
output channels split into three blocks -- where Owl's cumulative offsets
(:589-590, :854-855) go wrong and the restatement must not."""
import numpy as np
import pytest

import oracle
import oracle_ndarray as on
from owl_geometry import SAME


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("l1,fwd,bwd,one_by_one", [
    # 32 KiB L1: max_kc = ((32768 - 128) / 48) & ~7 = 680
    (32768, (576, 128, 340), (680, 576, 128), (680, 256, 256)),
    # 48 KiB L1: max_kc = 1016; the 1x1 256->512 layer's Co splits into blocks of 172
    (49152, (576, 128, 288), (1016, 576, 128), (1008, 256, 172)),
])
def test_block_sizes_known_answers(l1, fwd, bwd, one_by_one):
    # forward of the 64x56x56x64 -> 128 3x3 layer: (K, Co, M); backward: (M, K, Co)
    assert oracle.owl_block_sizes(576, 128, 200704, 4, l1, 1 << 20, 32 << 20) == fwd
    assert oracle.owl_block_sizes(200704, 576, 128, 4, l1, 1 << 20, 32 << 20) == bwd
    # backward of the 128x28x28x256 -> 512 1x1 stride-2 layer
    assert oracle.owl_block_sizes(25088, 256, 512, 4, l1, 1 << 20, 32 << 20) == one_by_one
    # all three below 50: unchanged
    assert oracle.owl_block_sizes(40, 30, 20, 4, l1, 1 << 20, 32 << 20) == (40, 30, 20)


def _layer(seed, xs, ks):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(xs).astype(np.float32)
    w = (rng.standard_normal(ks) / np.sqrt(np.prod(ks[:3]))).astype(np.float32)
    return x, w, rng


@pytest.mark.parametrize("cache", [(262144, 1 << 20, 8 << 20), None], ids=["co-in-3-blocks", "host-cache"])
def test_gebp_matches_fp64(cache):
    # M K = 4096 * 288 >= 512 Ki: the GEBP branch.  With a 256 KiB L1 the
    # backward passes take the whole M as one k block (4096 <= max_kc) and
    # split Co = 100 into blocks of 36, 36, 28.
    x, w, rng = _layer(1, [4, 32, 32, 32], [3, 3, 32, 100])
    if cache is not None:
        assert oracle.owl_block_sizes(4096, 288, 100, 4, *cache) == (4096, 288, 36)
    g = oracle.conv_geom(list(x.shape), list(w.shape), [32, 32], [1, 1], [1, 1], [1, 1])
    dy = rng.standard_normal([4, 32, 32, 100]).astype(np.float32)
    br = []
    y = oracle.conv_fwd_owl_f32(g, x, w, cache, br)
    dx = oracle.conv_bwd_input_owl_f32(g, w, dy, cache, br)
    dw = oracle.conv_bwd_kernel_owl_f32(g, x, dy, cache, br)
    assert br == ["gebp"] * 3
    x64, w64, dy64 = x.astype(np.float64), w.astype(np.float64), dy.astype(np.float64)
    assert rel(y.reshape(4, 32, 32, 100), on.conv2d(x64, w64, [1, 1], SAME)) < 1e-5
    assert rel(dx.reshape(x.shape), on.conv2d_backward_input(x64, w64, [1, 1], dy64)) < 1e-5
    assert rel(dw.reshape(w.shape), on.conv2d_backward_kernel(x64, w64, [1, 1], dy64)) < 1e-5


def test_gebp_on_benchmark_layer_slice():
    """Two images of the 64x56x56x64 -> 128 3x3 layer (M K = 6272 * 576: GEBP)
    with the host's cache sizes, as the CPU baseline runs it."""
    x, w, rng = _layer(2, [2, 56, 56, 64], [3, 3, 64, 128])
    g = oracle.conv_geom(list(x.shape), list(w.shape), [56, 56], [1, 1], [1, 1], [1, 1])
    dy = rng.standard_normal([2, 56, 56, 128]).astype(np.float32)
    br = []
    y = oracle.conv_fwd_owl_f32(g, x, w, None, br)
    dw = oracle.conv_bwd_kernel_owl_f32(g, x, dy, None, br)
    assert br == ["gebp", "gebp"]
    x64, w64, dy64 = x.astype(np.float64), w.astype(np.float64), dy.astype(np.float64)
    assert rel(y.reshape(2, 56, 56, 128), on.conv2d(x64, w64, [1, 1], SAME)) < 1e-5
    assert rel(dw.reshape(w.shape), on.conv2d_backward_kernel(x64, w64, [1, 1], dy64)) < 1e-5


def test_small_layers_take_im2col():
    x, w, rng = _layer(3, [2, 8, 8, 3], [3, 3, 3, 4])
    g = oracle.conv_geom(list(x.shape), list(w.shape), [8, 8], [1, 1], [1, 1], [1, 1])
    br = []
    y = oracle.conv_fwd_owl_f32(g, x, w, None, br)
    assert br == ["im2col"]
    assert rel(y.reshape(2, 8, 8, 4), on.conv2d(x.astype(np.float64), w.astype(np.float64), [1, 1], SAME)) < 1e-5
