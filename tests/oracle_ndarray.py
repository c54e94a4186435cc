"""Owl's Dense.Ndarray conv / pool functions evaluated by the CPU oracle.

Same signatures as ``owl_amd.ndarray`` (numpy in, numpy out).  Geometry
follows Owl's wrappers (owl_dense_ndarray_generic.ml) with the formulas of
``owl_geometry`` (restated apart from the product's ``owl_amd.shapes``); the
arithmetic is the oracle's FP64 direct loops (conv) or the bit-exact pooling
restatement in the input's own dtype.
Test infrastructure only.
"""
import numpy as np

import oracle
from owl_geometry import SAME, dilated_extent as k_eff, lead_pad as c_pad, out_dim as conv_out_dim, \
    transpose_dim as transpose_out_dim, wrapper_pads_2d as calc_conv2d_padding


def _geom(B, sp_in, Ci, ksp, Co, sp_out, stride, rate, pads):
    nd = len(sp_in)
    return oracle.conv_geom([B, *sp_in, Ci], [*ksp, Ci, Co], sp_out, stride, rate, pads) if nd == 3 else \
        oracle.conv_geom([B, *sp_in, Ci], [*ksp, Ci, Co], sp_out, stride, rate, pads)


def _pads(sp_in, ksp, sp_out, stride, rate):
    return [c_pad(s, o, k_eff(k, r), n) for n, k, o, s, r in zip(sp_in, ksp, sp_out, stride, rate)]


def _fwd(x, k, stride, rate, padding):
    nd = x.ndim - 2
    rate = list(rate) if rate is not None else [1] * nd
    B, sp_in, Ci = x.shape[0], list(x.shape[1:-1]), x.shape[-1]
    ksp, Co = list(k.shape[:-2]), k.shape[-1]
    sp_out = [conv_out_dim(padding, n, k_eff(kk, r), s) for n, kk, s, r in zip(sp_in, ksp, stride, rate)]
    pads = _pads(sp_in, ksp, sp_out, stride, rate) if padding == SAME else [0] * nd
    g = _geom(B, sp_in, Ci, ksp, Co, sp_out, stride, rate, pads)
    return oracle.conv_fwd(g, x, k).reshape([B, *sp_out, Co])


def _bwd(which, x, k, stride, rate, dy):
    nd = x.ndim - 2
    rate = list(rate) if rate is not None else [1] * nd
    B, sp_in, Ci = x.shape[0], list(x.shape[1:-1]), x.shape[-1]
    ksp, Co = list(k.shape[:-2]), k.shape[-1]
    sp_out = list(dy.shape[1:-1])
    g = _geom(B, sp_in, Ci, ksp, Co, sp_out, stride, rate, _pads(sp_in, ksp, sp_out, stride, rate))
    if which == "input":
        return oracle.conv_bwd_input(g, k, dy).reshape(x.shape)
    return oracle.conv_bwd_kernel(g, x, dy).reshape(k.shape)


def conv2d(x, k, stride, padding=SAME):
    return _fwd(x, k, stride, None, padding)


conv3d = conv2d


def conv2d_backward_input(x, k, stride, dy):
    return _bwd("input", x, k, stride, None, dy)


def conv2d_backward_kernel(x, k, stride, dy):
    return _bwd("kernel", x, k, stride, None, dy)


conv3d_backward_input = conv2d_backward_input
conv3d_backward_kernel = conv2d_backward_kernel


def dilated_conv2d(x, k, stride, rate, padding=SAME):
    return _fwd(x, k, stride, rate, padding)


dilated_conv3d = dilated_conv2d


def dilated_conv2d_backward_input(x, k, stride, rate, dy):
    return _bwd("input", x, k, stride, rate, dy)


def dilated_conv2d_backward_kernel(x, k, stride, rate, dy):
    return _bwd("kernel", x, k, stride, rate, dy)


dilated_conv3d_backward_input = dilated_conv2d_backward_input
dilated_conv3d_backward_kernel = dilated_conv2d_backward_kernel


def _tgeom(x, k, sp_out, stride):
    nd = x.ndim - 2
    B, sp_in, Ci = x.shape[0], list(x.shape[1:-1]), x.shape[-1]
    ksp, Co = list(k.shape[:-2]), k.shape[-1]
    pads = _pads(sp_out, ksp, sp_in, stride, [1] * nd)
    return oracle.conv_geom([B, *sp_out, Co], [*ksp, Co, Ci], sp_in, stride, [1] * nd, pads)


def transpose_conv2d(x, k, stride, padding=SAME):
    nd = x.ndim - 2
    sp_out = [transpose_out_dim(padding, n, kk, s) for n, kk, s in zip(x.shape[1:-1], k.shape[:-2], stride)]
    g = _tgeom(x, k, sp_out, stride)
    return oracle.conv_bwd_input(g, k, x).reshape([x.shape[0], *sp_out, k.shape[-1]])


transpose_conv3d = transpose_conv2d


def transpose_conv2d_backward_input(x, k, stride, dy):
    g = _tgeom(x, k, list(dy.shape[1:-1]), stride)
    return oracle.conv_fwd(g, dy, k).reshape(x.shape)


def transpose_conv2d_backward_kernel(x, k, stride, dy):
    g = _tgeom(x, k, list(dy.shape[1:-1]), stride)
    return oracle.conv_bwd_kernel(g, dy, x).reshape(k.shape)


transpose_conv3d_backward_input = transpose_conv2d_backward_input
transpose_conv3d_backward_kernel = transpose_conv2d_backward_kernel


# ------------------------------------------------------------------ pooling
def _pgeom(x, kernel, stride, padding):
    sp_in = list(x.shape[1:-1])
    sp_out = [conv_out_dim(padding, n, kk, s) for n, kk, s in zip(sp_in, kernel, stride)]
    if len(sp_in) == 2:
        top, left, _, _ = calc_conv2d_padding(sp_in[0], sp_in[1], kernel[0], kernel[1], sp_out[0], sp_out[1],
                                              stride[1], stride[0])
        pads = [left, top] if padding == SAME else [0, 0]
    else:
        pads = [c_pad(s, o, kk, n) for n, kk, o, s in zip(sp_in, kernel, sp_out, stride)] \
            if padding == SAME else [0, 0, 0]
    return oracle.pool_geom(list(x.shape), kernel, sp_out, stride, pads), sp_out


def _pool(x, kernel, stride, padding, is_max):
    g, sp_out = _pgeom(x, kernel, stride, padding)
    return oracle.pool_fwd(g, x, is_max).reshape([x.shape[0], *sp_out, x.shape[-1]])


def max_pool2d(x, kernel, stride, padding=SAME):
    return _pool(x, kernel, stride, padding, True)


def avg_pool2d(x, kernel, stride, padding=SAME):
    return _pool(x, kernel, stride, padding, False)


max_pool3d, avg_pool3d = max_pool2d, avg_pool2d


def max_pool2d_argmax(x, kernel, stride, padding=SAME):
    g, sp_out = _pgeom(x, kernel, stride, padding)
    y, a = oracle.pool_argmax(g, x)
    shp = [x.shape[0], *sp_out, x.shape[-1]]
    return y.reshape(shp), a.reshape(shp)


def _pool_bwd(padding, x, kernel, stride, dy, is_max):
    g, _ = _pgeom(x, kernel, stride, padding)
    return oracle.pool_bwd(g, x, dy, is_max).reshape(x.shape)


def max_pool2d_backward(padding, x, kernel, stride, dy):
    return _pool_bwd(padding, x, kernel, stride, dy, True)


def avg_pool2d_backward(padding, x, kernel, stride, dy):
    return _pool_bwd(padding, x, kernel, stride, dy, False)


max_pool3d_backward, avg_pool3d_backward = max_pool2d_backward, avg_pool2d_backward
