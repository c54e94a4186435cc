"""Output-shape and padding arithmetic of Owl's conv / pool wrappers,
restated for the test side only, independently of ``owl_amd.shapes``.

Owl computes the output sizes in floats (``owl_utils_infer_shape.ml:19-44,
127-163``: SAME ``ceil(n / s)``, VALID ``ceil((n - k + 1) / s)``, then
``int_of_float``).  Here the same values come from integer arithmetic:
``ceil(a / s) == -((-a) // s)`` for any integer ``a`` and ``s > 0`` (Python's
floor division), which is exact where the float form is, and equals
``int_of_float (ceil x)`` for negative ``x`` too.  Transposed sizes follow
``owl_utils_infer_shape.ml:48-67``; the wrapper pads ``:71-91``; the C side's
leading pad ``(s (o - 1) + k_eff - n) / 2`` (C truncation, clamped at zero,
``owl_ndarray_conv_impl.h:288-294, 502-505, 713-716, 3144-3150``) is
``max(t, 0) // 2``, the clamp making the truncation and the floor agree.

The oracle mirror (``oracle_ndarray``) and the oracle identity tests use this
module; ``test_oracle_identities.test_geometry_matches_product`` checks the
product's ``owl_amd.shapes`` against it over a brute-force grid.
"""

SAME = "SAME"
VALID = "VALID"


def _ceil_div(a, s):
    return -((-a) // s)


def out_dim(padding, n, k, s):
    """Output extent of one axis (SAME / VALID)."""
    if padding == SAME:
        return _ceil_div(n, s)
    if padding == VALID:
        return _ceil_div(n - k + 1, s)
    raise ValueError(padding)


def transpose_dim(padding, n, k, s):
    """Output extent of one axis of a transposed convolution."""
    if padding == SAME:
        return n * s
    if padding == VALID:
        return n * s + (k - s if k > s else 0)
    raise ValueError(padding)


def dilated_extent(k, r):
    """Extent of a kernel axis of k taps at rate r."""
    return (k - 1) * r + 1


def lead_pad(s, o, ke, n):
    """Leading pad the C side derives from the shapes."""
    t = s * (o - 1) + ke - n
    return (t if t > 0 else 0) // 2


def wrapper_pads_2d(n_cols, n_rows, k_cols, k_rows, o_cols, o_rows, row_stride, col_stride):
    """(top, left, bottom, right) as Owl's wrappers pass them to the 2-D
    pooling backward / argmax stubs: "top" along rows (W), "left" along
    cols (H)."""
    along_h = (o_rows - 1) * row_stride + k_rows - n_rows
    along_w = (o_cols - 1) * col_stride + k_cols - n_cols
    along_h, along_w = max(along_h, 0), max(along_w, 0)
    return along_h // 2, along_w // 2, along_h - along_h // 2, along_w - along_w // 2
