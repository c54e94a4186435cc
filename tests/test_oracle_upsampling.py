"""The oracle's restatement of Owl's upsampling (owl_ndarray_upsampling_impl.h
for the backward, ``repeat`` for the forward) against the reference's own
test cases, test/unit_upsampling_generic.ml:

* forward (fun00-fun03): ``upsampling2d (sequential shape) size`` equals
  ``repeat x [|1; s0; s1; 1|]``, checked against numpy's ``repeat`` (a
  closed form independent of the oracle);
* backward (fun04-fun08): ``upsampling2d_backward (ones shape) size
  (sequential out_shape)`` equals the transcribed expected values.

All values are small integers, so FP32 and FP64 must match exactly."""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "owl_upsampling_vectors.json")))["cases"]
FWD = [c for c in CASES if c["helper"] == "test_upsampling2d_forward"]
BWD = [c for c in CASES if c["helper"] == "upsampling2d_backward_verify_value"]


def sequential(shape, dt):
    return np.arange(int(np.prod(shape)), dtype=dt).reshape(shape)


def test_fixture_covers_the_reference_tests():
    assert len(CASES) == 9 and all(c["registered"] for c in CASES)
    assert len(FWD) == 4 and len(BWD) == 5


@pytest.mark.parametrize("case", FWD, ids=[c["fun"] for c in FWD])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_forward_is_repeat(case, dt):
    x = sequential(case["input_shape"], dt)
    s0, s1 = case["size"]
    y = oracle.upsampling_fwd(x, case["size"])
    assert np.array_equal(y, x.repeat(s0, axis=1).repeat(s1, axis=2))


@pytest.mark.parametrize("case", BWD, ids=[c["fun"] for c in BWD])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_backward_golden(case, dt):
    shp = case["input_shape"]
    s0, s1 = case["size"]
    dy = sequential([shp[0], shp[1] * s0, shp[2] * s1, shp[3]], dt)
    dx = oracle.upsampling_bwd(dy, shp, case["size"])
    assert np.array_equal(dx.ravel(), np.array(case["expected"], dt))


def test_backward_is_adjoint_of_forward():
    # <up(x), g> = <x, up_bwd(g)> with the forward's exact extent
    rng = np.random.default_rng(0)
    x = rng.standard_normal([2, 3, 5, 4])
    g = rng.standard_normal([2, 6, 15, 4])
    lhs = np.vdot(oracle.upsampling_fwd(x, [2, 3]), g)
    rhs = np.vdot(x, oracle.upsampling_bwd(g, x.shape, [2, 3]))
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_backward_clamps_like_owl():
    # outputs past the scaled extent land on the last input row / column
    # (5 output rows at scale 2: rows 0-1 -> 0, rows 2-4 -> 1; 4 output
    # columns at scale 1: column 0 -> 0, columns 1-3 -> 1)
    dy = np.ones([1, 5, 4, 1])
    dx = oracle.upsampling_bwd(dy, [1, 2, 2, 1], [2, 1])
    assert np.array_equal(dx[0, :, :, 0], np.array([[2.0, 6.0], [3.0, 9.0]]))
