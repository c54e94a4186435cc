"""The CPU oracle against Owl's own golden vectors (test/unit_*conv*.ml,
test/unit_pool*.ml), with Owl's acceptance rule sum|a - b| < 5e-4
(test/unit_conv2d_generic.ml:12-14) and exact equality where every expected
value is an integer."""
import numpy as np
import pytest

import oracle_ndarray
import owl_cases

CASES = owl_cases.load_cases()


@pytest.mark.parametrize("c", CASES, ids=[owl_cases.case_id(c) for c in CASES])
def test_oracle_matches_owl_golden(c):
    r = owl_cases.run_case(c, oracle_ndarray, np.float32 if "pool" in c["file"] else np.float64)
    e = owl_cases.expected_of(c, r.shape)
    assert np.abs(r.astype(np.float64) - e).sum() < 5e-4
    if np.all(e == np.round(e)) and "pool" not in c["file"]:
        # integer-valued fixture: the FP64 oracle is exact
        assert np.array_equal(r.astype(np.float64), e)


def test_fixture_covers_every_family():
    files = {c["file"] for c in CASES}
    assert len(CASES) >= 190
    for f in ("unit_conv2d_generic", "unit_conv3d_generic", "unit_pool2d_generic", "unit_pool3d_generic",
              "unit_dilated_conv2d", "unit_dilated_conv3d", "unit_transpose_conv2d", "unit_transpose_conv3d"):
        assert f"test/{f}.ml" in files
