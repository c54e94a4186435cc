"""Output-shape and padding formulas of Owl's conv / pool wrappers.

Restates ``src/base/misc/owl_utils_infer_shape.ml`` exactly, including Owl's
axis naming: for an NHWC array ``[b; h; w; c]`` Owl calls ``h`` "cols" and
``w`` "rows"; ``stride.(0)`` is the column (H) stride and ``stride.(1)`` the
row (W) stride (``owl_dense_ndarray_generic.ml:2358-2361``).
"""
import math

SAME = "SAME"
VALID = "VALID"


class InvalidArgument(ValueError):
    """Mirror of ``Owl_exception.INVALID_ARGUMENT`` raised by the wrappers."""


def _pad_code(padding):
    if padding == SAME:
        return 0
    if padding == VALID:
        return 1
    raise InvalidArgument(f"padding must be SAME or VALID, got {padding!r}")


def conv_out_dim(padding, n, k, s):
    """One axis of ``calc_conv2d_output_shape`` (owl_utils_infer_shape.ml:19-44).

    Owl computes in floats: SAME ``ceil(n / s)``, VALID ``ceil((n - k + 1) / s)``.
    ``int_of_float`` truncates, which matters only for a non-positive VALID size.
    """
    if padding == SAME:
        return int(math.ceil(float(n) / float(s)))
    if padding == VALID:
        return int(math.ceil((float(n) - float(k) + 1.0) / float(s)))
    raise InvalidArgument(f"padding must be SAME or VALID, got {padding!r}")


def calc_conv2d_output_shape(padding, input_cols, input_rows, kernel_cols, kernel_rows,
                             row_stride, col_stride):
    """owl_utils_infer_shape.ml:19-44."""
    return (conv_out_dim(padding, input_cols, kernel_cols, col_stride),
            conv_out_dim(padding, input_rows, kernel_rows, row_stride))


def calc_conv3d_output_shape(padding, input_cols, input_rows, input_dpts, kernel_cols,
                             kernel_rows, kernel_dpts, row_stride, col_stride, dpt_stride):
    """owl_utils_infer_shape.ml:127-163."""
    return (conv_out_dim(padding, input_cols, kernel_cols, col_stride),
            conv_out_dim(padding, input_rows, kernel_rows, row_stride),
            conv_out_dim(padding, input_dpts, kernel_dpts, dpt_stride))


def transpose_out_dim(padding, n, k, s):
    """One axis of ``calc_transpose_conv2d_output_shape`` (owl_utils_infer_shape.ml:48-67)."""
    if padding == SAME:
        return n * s
    if padding == VALID:
        return n * s + max(k - s, 0)
    raise InvalidArgument(f"padding must be SAME or VALID, got {padding!r}")


def calc_conv2d_padding(input_cols, input_rows, kernel_cols, kernel_rows, output_cols,
                        output_rows, row_stride, col_stride):
    """owl_utils_infer_shape.ml:71-91 -> (pad_top, pad_left, pad_bottom, pad_right).

    Owl's "top" is computed along rows (W) and "left" along cols (H); the C
    pooling kernels receive them as ``pad_rows`` and ``pad_cols``.
    """
    ph = max((output_rows - 1) * row_stride + kernel_rows - input_rows, 0)
    pw = max((output_cols - 1) * col_stride + kernel_cols - input_cols, 0)
    top, left = ph // 2, pw // 2
    return top, left, ph - top, pw - left


def c_pad(stride, out, k_eff, n):
    """Leading pad as Owl's C derives it from the shapes.

    ``p = (s*(o-1) + k_eff - n) / 2`` with C's truncating division, clamped at
    zero (owl_ndarray_conv_impl.h:288-294, 502-505, 713-716, 3144-3150).
    """
    t = stride * (out - 1) + k_eff - n
    p = int(t / 2) if t >= 0 else -((-t) // 2)
    return max(p, 0)


def k_eff(k, rate):
    """Dilated kernel extent ``k + (k-1)(r-1)`` (owl_ndarray_conv_impl.h:3141-3142)."""
    return k + (k - 1) * (rate - 1)
