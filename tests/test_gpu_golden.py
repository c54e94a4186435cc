"""Owl's golden vectors through the Owl drop-in entry points on the GPU
(numpy host buffers in, Owl's stub argument order, synchronous), for Owl's
S (float32) and D (float64) modules."""
import numpy as np
import pytest

import oracle_ndarray
import owl_cases
from owl_amd import ndarray as nd

CASES = owl_cases.load_cases()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["S", "D"])
@pytest.mark.parametrize("c", CASES, ids=[owl_cases.case_id(c) for c in CASES])
def test_gpu_dropin_matches_owl_golden(c, dtype):
    r = owl_cases.run_case(c, nd, dtype)
    assert r.dtype == dtype
    e = owl_cases.expected_of(c, r.shape)
    assert np.abs(r.astype(np.float64) - e).sum() < 5e-4
    if "pool" in c["file"]:
        # pooling is bit-identical to Owl's C (restated by the oracle)
        ref = owl_cases.run_case(c, oracle_ndarray, dtype)
        assert np.array_equal(r, ref)
    elif np.all(e == np.round(e)) and np.abs(e).max() < 2 ** 24:
        # integer-valued products and sums are exact in FP32 and FP64
        assert np.array_equal(r.astype(np.float64), e)
