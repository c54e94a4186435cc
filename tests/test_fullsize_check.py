"""The full-size weight-gradient sampler against the FP64 oracle on small
problems (every entry), and what it must reject: an all-zero gradient, one
with a transposed tap, one from the wrong pad."""
import numpy as np
import pytest

import oracle_ndarray as on
from fullsize import dw_entries, pick_images, sample_dw_index, sampled_rel
from owl_geometry import SAME, VALID


@pytest.mark.parametrize("seed", range(6))
def test_sampler_matches_oracle_on_every_entry(seed):
    rng = np.random.default_rng(seed)
    B, H, W = int(rng.integers(1, 4)), int(rng.integers(4, 10)), int(rng.integers(4, 10))
    Ci, Co = int(rng.integers(1, 5)), int(rng.integers(1, 5))
    k = int(rng.integers(1, 4))
    stride = [int(rng.integers(1, 3)), int(rng.integers(1, 3))]
    pad = [SAME, VALID][seed % 2]
    x = rng.standard_normal([B, H, W, Ci])
    w = rng.standard_normal([k, k, Ci, Co])
    y = on.conv2d(x, w, stride, pad)
    dy = rng.standard_normal(y.shape)
    ref = on.conv2d_backward_kernel(x, w, stride, dy)
    idx = sample_dw_index([k, k], Ci, Co, 10 ** 6, seed)
    assert len(idx) == ref.size
    got = dw_entries(x, dy, [k, k], stride, idx)
    np.testing.assert_allclose(got, ref[tuple(idx.T)], rtol=1e-12, atol=1e-12)


def test_sampler_rejects_wrong_gradients():
    rng = np.random.default_rng(7)
    x = rng.standard_normal([3, 12, 11, 4])
    w = rng.standard_normal([3, 3, 4, 5])
    dy = rng.standard_normal([3, 12, 11, 5])
    ref = on.conv2d_backward_kernel(x, w, [1, 1], dy)
    idx = sample_dw_index([3, 3], 4, 5, 64, 1)
    samples = dw_entries(x, dy, [3, 3], [1, 1], idx)
    assert sampled_rel(ref[tuple(idx.T)], samples) < 1e-12
    assert sampled_rel(np.zeros_like(ref)[tuple(idx.T)], samples) == 1.0
    flipped = ref[::-1, ::-1]                      # taps in the wrong order
    assert sampled_rel(flipped[tuple(idx.T)], samples) > 0.1
    shifted = on.conv2d_backward_kernel(x[:, 1:], w, [1, 1], dy[:, :-1])  # one row off
    assert sampled_rel(shifted[tuple(idx.T)], samples) > 0.1


def test_pick_images_skips_the_front():
    im = pick_images(64, 2, 3)
    assert len(im) == 2 and len(set(im)) == 2 and min(im) >= 2 and max(im) < 64
