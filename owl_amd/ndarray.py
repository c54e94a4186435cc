"""Owl's ``Dense.Ndarray`` convolution / pooling API on the MI355X path.

Each function mirrors the OCaml wrapper of the same name in
``src/owl/dense/owl_dense_ndarray_generic.ml`` (argument order, shape checks,
output shape, SAME/VALID handling, stride/rate unpacking) and then calls the
C ABI of ``libowl_amd.so``:

* numpy arrays (float32 = Owl's S, float64 = Owl's D) go through the Owl
  drop-in entry points (``owl_amd_<dtype>_ndarray_*``) with Owl's stub
  arguments, exactly as the OCaml externals would call them;
* torch CUDA tensors (float32, float64, bfloat16, float16) stay in HBM and go
  through the device entry points on torch's current stream.

Shapes use Owl's order: input ``[b; h; w; c]`` (3-D ``[b; h; w; d; c]``),
kernel ``[kh; kw; ci; co]``, ``stride = [| h_stride; w_stride |]``.
"""
import ctypes

import numpy as np

from . import _lib
from .shapes import (SAME, VALID, InvalidArgument, c_pad, calc_conv2d_padding, conv_out_dim,
                     k_eff, transpose_out_dim)

try:  # torch is plumbing for device memory and streams only
    import torch
except Exception:  # pragma: no cover
    torch = None

__all__ = [
    "conv2d", "conv2d_backward_input", "conv2d_backward_kernel",
    "conv3d", "conv3d_backward_input", "conv3d_backward_kernel",
    "dilated_conv2d", "dilated_conv2d_backward_input", "dilated_conv2d_backward_kernel",
    "dilated_conv3d", "dilated_conv3d_backward_input", "dilated_conv3d_backward_kernel",
    "transpose_conv2d", "transpose_conv2d_backward_input", "transpose_conv2d_backward_kernel",
    "transpose_conv3d", "transpose_conv3d_backward_input", "transpose_conv3d_backward_kernel",
    "max_pool2d", "avg_pool2d", "max_pool3d", "avg_pool3d", "max_pool2d_argmax",
    "max_pool2d_backward", "avg_pool2d_backward", "max_pool3d_backward", "avg_pool3d_backward",
    "SAME", "VALID", "InvalidArgument",
]


# ----------------------------------------------------------------------------- helpers
def _is_torch(a):
    return torch is not None and isinstance(a, torch.Tensor)


def _np_prefix(a):
    if a.dtype == np.float32:
        return "float32"
    if a.dtype == np.float64:
        return "float64"
    if a.dtype == np.complex64:
        return "complex32"   # Owl's C: pairs of float32
    if a.dtype == np.complex128:
        return "complex64"   # Owl's Z: pairs of float64
    raise InvalidArgument(f"unsupported dtype {a.dtype} (Owl S, D, C, Z = float32, float64, complex64, complex128)")


_TORCH_DT = {}
if torch is not None:
    _TORCH_DT = {torch.float32: _lib.F32, torch.float64: _lib.F64,
                 torch.bfloat16: _lib.BF16, torch.float16: _lib.F16,
                 torch.complex64: _lib.C32, torch.complex128: _lib.C64}


def _tdt(t):
    if t.dtype not in _TORCH_DT:
        raise InvalidArgument(f"unsupported tensor dtype {t.dtype}")
    return _TORCH_DT[t.dtype]


def _ptr(a):
    if _is_torch(a):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _contig(a):
    if _is_torch(a):
        if not a.is_cuda:
            raise InvalidArgument("torch tensors must be on a CUDA (HIP) device")
        return a.contiguous()
    return np.ascontiguousarray(a)


def _empty_like_kind(ref, shape, dtype=None):
    if _is_torch(ref):
        return torch.empty(shape, dtype=dtype or ref.dtype, device=ref.device)
    return np.empty(shape, dtype=dtype or ref.dtype)


def _verify(cond, msg):
    if not cond:
        raise InvalidArgument(msg)


def _overlaps(a, b):
    if _is_torch(a):
        a0, b0 = a.data_ptr(), b.data_ptr()
        return a0 < b0 + b.numel() * b.element_size() and b0 < a0 + a.numel() * a.element_size()
    return np.may_share_memory(a, b)


def _out_buf(out, ref, shape, dtype=None, name="", inputs=()):
    """The result array: a new one, or the caller's ``out`` (Owl's ``*_ ~out``
    variants write into a given array) after checking that it is the same kind
    of array as the input, has the result's shape and dtype, is contiguous and
    writeable, and does not overlap an input."""
    if out is None:
        return _empty_like_kind(ref, shape, dtype=dtype)
    want = dtype or ref.dtype
    _verify(_is_torch(out) == _is_torch(ref), f"{name}: out is not the same kind of array as the input")
    _verify(list(out.shape) == list(shape), f"{name}: out shape {list(out.shape)} should be {list(shape)}")
    _verify(out.dtype == want, f"{name}: out dtype {out.dtype} should be {want}")
    if _is_torch(out):
        _verify(out.is_contiguous() and out.device == ref.device, f"{name}: out must be contiguous, on the input's device")
    else:
        _verify(out.flags.c_contiguous and out.flags.writeable, f"{name}: out must be a writeable C-contiguous array")
    for a in inputs:
        _verify(not _overlaps(out, a), f"{name}: out overlaps an input")
    return out


def _same_kind(*arrs):
    t = [_is_torch(a) for a in arrs]
    _verify(all(t) or not any(t), "mix of numpy arrays and device tensors")
    if not t[0]:
        _verify(len({a.dtype for a in arrs}) == 1, "arguments must share one dtype")


def _desc(B, H, W, D, Ci, kh, kw, kd, Co, Ho, Wo, Do, s, r, pads):
    return _lib.ConvDesc(B, H, W, D, Ci, kh, kw, kd, Co, Ho, Wo, Do, *s, *r, *pads)


# ----------------------------------------------------------------------------- conv core
# A conv problem in the library's 3-D form; 2-D problems have D = kd = Do = 1.
class _Conv:
    def __init__(self, B, sp_in, Ci, ksp, Co, sp_out, stride, rate, pads):
        self.B, self.sp_in, self.Ci, self.ksp, self.Co = B, list(sp_in), Ci, list(ksp), Co
        self.sp_out, self.stride, self.rate, self.pads = list(sp_out), list(stride), list(rate), list(pads)

    def desc(self):
        f = lambda v, d: list(v) + [d] * (3 - len(v))
        return _desc(self.B, *f(self.sp_in, 1), self.Ci, *f(self.ksp, 1), self.Co,
                     *f(self.sp_out, 1), f(self.stride, 1), f(self.rate, 1), f(self.pads, 0))


def _pads_from_shapes(sp_in, ksp, sp_out, stride, rate):
    return [c_pad(s, o, k_eff(k, r), n) for n, k, o, s, r in zip(sp_in, ksp, sp_out, stride, rate)]


def conv_problem(x_shape, k_shape, stride, rate=None, padding=SAME):
    """Geometry (a ``_Conv``) of ``dilated_conv{2,3}d ~padding x w stride rate``
    exactly as the forward wrappers derive it."""
    nd_ = len(x_shape) - 2
    rate = list(rate) if rate is not None else [1] * nd_
    B, sp_in, Ci = x_shape[0], list(x_shape[1:-1]), x_shape[-1]
    ksp, Co = list(k_shape[:-2]), k_shape[-1]
    _verify(Ci == k_shape[-2], "input channels must equal the kernel's third dimension")
    sp_out = [conv_out_dim(padding, n, k_eff(kk, r), s) for n, kk, s, r in zip(sp_in, ksp, stride, rate)]
    _verify(all(o > 0 for o in sp_out), f"empty output shape {sp_out}")
    pads = _pads_from_shapes(sp_in, ksp, sp_out, stride, rate) if padding == SAME else [0] * nd_
    return _Conv(B, sp_in, Ci, ksp, Co, sp_out, stride, rate, pads)


def _dev_conv(pass_, g, a, b, out, out_t=None, workspace=None):
    """Run one pass on device tensors.  pass_: fwd(a=x, b=w), bi(a=w, b=dy), bk(a=x, b=dy).
    ``workspace``: a uint8 device tensor for the backward-kernel pass's
    split-K partials (``_ws_bytes``), else the library's own per-device one."""
    st = _stream()
    d = g.desc()
    it = _tdt(a)
    ot = _TORCH_DT[out.dtype] if out_t is None else out_t
    L = _lib.lib()
    if pass_ == "fwd":
        _lib.check(L.owl_amd_conv_fwd(ctypes.byref(d), it, ot, _ptr(a), _ptr(b), _ptr(out), st))
    elif pass_ == "bi":
        _lib.check(L.owl_amd_conv_bwd_input(ctypes.byref(d), it, ot, _ptr(a), _ptr(b), _ptr(out), st))
    else:
        ws, nb = (None, 0) if workspace is None else (_ptr(workspace), workspace.numel())
        _lib.check(L.owl_amd_conv_bwd_kernel(ctypes.byref(d), it, ot, _ptr(a), _ptr(b), _ptr(out),
                                            ws, nb, st))
    return out


def _host_conv2d(variant, x, w, y, g, padding_code=None, rate=(1, 1)):
    """Call an Owl 2-D conv drop-in with Owl's stub argument order
    (owl_ndarray_conv_impl.h:246-253 / :465-471)."""
    H, W = g.sp_in
    kh, kw = g.ksp
    Ho, Wo = g.sp_out
    sh, sw = g.stride
    args = [_ptr(x), _ptr(w), _ptr(y), g.B, H, W, g.Ci, kh, kw, Ho, Wo, g.Co, sw, sh]
    if padding_code is not None:
        args.append(padding_code)
    args += [rate[1], rate[0]]
    _lib.call(f"owl_amd_{_np_prefix(x)}_ndarray_conv_{variant}", *args)


def _host_conv3d(variant, x, w, y, g, padding_code=None, rate=None):
    """Owl 3-D stub order (owl_ndarray_conv_impl.h:1211-1220, :3417-3427)."""
    H, W, D = g.sp_in
    kh, kw, kd = g.ksp
    Ho, Wo, Do = g.sp_out
    sh, sw, sd = g.stride
    args = [_ptr(x), _ptr(w), _ptr(y), g.B, H, W, D, g.Ci, kh, kw, kd, Ho, Wo, Do, g.Co, sd, sw, sh]
    if rate is not None:
        args += [rate[2], rate[1], rate[0]]
    if padding_code is not None:
        args.append(padding_code)
    _lib.call(f"owl_amd_{_np_prefix(x)}_ndarray_conv_{variant}", *args)


def _check_conv_args(name, nd, x, k, stride, rate=None, dy=None):
    _verify(len(x.shape) == nd + 2, f"{name}: input dimension = {len(x.shape)} (should be {nd + 2})")
    _verify(len(k.shape) == nd + 2, f"{name}: kernel dimension = {len(k.shape)} (should be {nd + 2})")
    _verify(len(stride) == nd, f"{name}: stride dimension = {len(stride)} (should be {nd})")
    if rate is not None:
        _verify(len(rate) == nd, f"{name}: rate dimension = {len(rate)} (should be {nd})")
    _verify(x.shape[-1] == k.shape[-2],
            f"{name}: the {nd + 2}th dimension of input shape should be equal to the "
            f"{nd + 1}rd dimension of kernel shape")
    if dy is not None:
        _verify(len(dy.shape) == nd + 2, f"{name}: output' dimension = {len(dy.shape)} (should be {nd + 2})")
        _verify(dy.shape[0] == x.shape[0] and dy.shape[-1] == k.shape[-1],
                f"{name}: output' shape does not match input / kernel")


def _conv_fwd(name, nd, x, k, stride, rate, padding, variant, out_dtype=None, out=None):
    _same_kind(x, k)
    _check_conv_args(name, nd, x, k, stride, rate)
    x, k = _contig(x), _contig(k)
    B, sp_in, Ci = x.shape[0], list(x.shape[1:-1]), x.shape[-1]
    ksp, Co = list(k.shape[:-2]), k.shape[-1]
    rate = list(rate) if rate is not None else [1] * nd
    sp_out = [conv_out_dim(padding, n, k_eff(kk, r), s) for n, kk, s, r in zip(sp_in, ksp, stride, rate)]
    _verify(all(o > 0 for o in sp_out), f"{name}: empty output shape {sp_out}")
    pads = _pads_from_shapes(sp_in, ksp, sp_out, stride, rate) if padding == SAME else [0] * nd
    g = _Conv(B, sp_in, Ci, ksp, Co, sp_out, stride, rate, pads)
    y = _out_buf(out, x, [B, *sp_out, Co], out_dtype, name, (x, k))
    if _is_torch(x):
        return _dev_conv("fwd", g, x, k, y)
    pc = 0 if padding == SAME else 1
    if nd == 2:
        _host_conv2d(variant, x, k, y, g, pc, rate)
    else:
        _host_conv3d(variant, x, k, y, g, pc, rate if variant.startswith("dilated") else None)
    return y


def _ws_bytes(g, t):
    """Bytes of split-K workspace the backward-kernel pass of geometry g needs for input type t."""
    return int(_lib.lib().owl_amd_conv_bwd_kernel_workspace(ctypes.byref(g.desc()), _tdt(t)))


def _bwd_geom(nd, xs, ks, stride, rate, dys):
    """Geometry of a (dilated) conv's backward passes from the shapes of x, w, dy."""
    rate = list(rate) if rate is not None else [1] * nd
    sp_in, ksp, sp_out = list(xs[1:-1]), list(ks[:-2]), list(dys[1:-1])
    pads = _pads_from_shapes(sp_in, ksp, sp_out, stride, rate)
    return _Conv(xs[0], sp_in, xs[-1], ksp, ks[-1], sp_out, stride, rate, pads)


def _tbwd_geom(nd, xs, ks, stride, dys):
    """Geometry of a transposed conv's backward passes: the conv whose input is
    dy ([.., Co]) and whose output is x ([.., Ci])."""
    sp_in, ksp, sp_out = list(xs[1:-1]), list(ks[:-2]), list(dys[1:-1])
    pads = _pads_from_shapes(sp_out, ksp, sp_in, stride, [1] * nd)
    return _Conv(xs[0], sp_out, ks[-1], ksp, xs[-1], sp_in, stride, [1] * nd, pads)


def _conv_bwd(name, nd, which, x, k, stride, rate, dy, variant, out_dtype=None, out=None, workspace=None):
    _same_kind(x, k, dy)
    _check_conv_args(name, nd, x, k, stride, rate, dy)
    x, k, dy = _contig(x), _contig(k), _contig(dy)
    rate = list(rate) if rate is not None else [1] * nd
    g = _bwd_geom(nd, list(x.shape), list(k.shape), stride, rate, list(dy.shape))
    if which == "input":
        res = _out_buf(out, x, list(x.shape), out_dtype, name, (x, k, dy))
        if _is_torch(x):
            return _dev_conv("bi", g, k, dy, res)
        args = (res, k, dy)
    else:
        res = _out_buf(out, k, list(k.shape), out_dtype, name, (x, k, dy))
        if _is_torch(x):
            return _dev_conv("bk", g, x, dy, res, workspace=workspace)
        args = (x, res, dy)
    if nd == 2:
        _host_conv2d(variant, *args, g, None, rate)
    else:
        _host_conv3d(variant, *args, g, None, rate if variant.startswith("dilated") else None)
    return res


# ----------------------------------------------------------------------------- conv2d / 3d
def conv2d(input, kernel, stride, padding=SAME, out_dtype=None, out=None):
    """owl_dense_ndarray_generic.ml:2324-2400."""
    return _conv_fwd("conv2d", 2, input, kernel, stride, None, padding, "spatial", out_dtype, out=out)


def conv2d_backward_input(input, kernel, stride, output_grad, out_dtype=None, out=None):
    """owl_dense_ndarray_generic.ml:2479-2560."""
    return _conv_bwd("conv2d_backward_input", 2, "input", input, kernel, stride, None, output_grad,
                     "spatial_backward_input", out_dtype, out=out)


def conv2d_backward_kernel(input, kernel, stride, output_grad, out_dtype=None, out=None):
    """owl_dense_ndarray_generic.ml:2660-2740."""
    return _conv_bwd("conv2d_backward_kernel", 2, "kernel", input, kernel, stride, None, output_grad,
                     "spatial_backward_kernel", out_dtype, out=out)


def _conv_bwd_both(name, nd, x, k, stride, rate, dy, dx_dtype=None, dw_dtype=None, workspace=None):
    _verify(_is_torch(x) and _is_torch(k) and _is_torch(dy), f"{name}: the paired backward takes device tensors")
    _same_kind(x, k, dy)
    _check_conv_args(name, nd, x, k, stride, rate, dy)
    x, k, dy = _contig(x), _contig(k), _contig(dy)
    rate = list(rate) if rate is not None else [1] * nd
    g = _bwd_geom(nd, list(x.shape), list(k.shape), stride, rate, list(dy.shape))
    dx = _out_buf(None, x, list(x.shape), dx_dtype, name, (x, k, dy))
    dw = _out_buf(None, k, list(k.shape), dw_dtype, name, (x, k, dy))
    ws, nb = (None, 0) if workspace is None else (_ptr(workspace), workspace.numel())
    d = g.desc()
    _lib.check(_lib.lib().owl_amd_conv_bwd(ctypes.byref(d), _tdt(x), _TORCH_DT[dx.dtype], _TORCH_DT[dw.dtype],
                                           _ptr(x), _ptr(k), _ptr(dy), _ptr(dx), _ptr(dw), ws, nb, _stream()))
    return dx, dw


def conv2d_backward(input, kernel, stride, output_grad, out_dtype=None, kernel_out_dtype=None, workspace=None):
    """Both backward passes of conv2d: (conv2d_backward_input,
    conv2d_backward_kernel), the pair Owl's reverse rule for conv2d computes
    (owl_algodiff_ops.ml:2364), in one library call on the current stream.
    Device tensors; out_dtype is dx's type, kernel_out_dtype dw's."""
    return _conv_bwd_both("conv2d_backward", 2, input, kernel, stride, None, output_grad, out_dtype,
                          kernel_out_dtype, workspace)


def conv3d_backward(input, kernel, stride, output_grad, out_dtype=None, kernel_out_dtype=None, workspace=None):
    """Both backward passes of conv3d (owl_algodiff_ops.ml's conv3d reverse
    rule), as ``conv2d_backward``."""
    return _conv_bwd_both("conv3d_backward", 3, input, kernel, stride, None, output_grad, out_dtype,
                          kernel_out_dtype, workspace)


def conv3d(input, kernel, stride, padding=SAME, out_dtype=None, out=None):
    """owl_dense_ndarray_generic.ml:2848 (dispatches to cuboid_im2col, owl_ndarray_conv.ml:826)."""
    return _conv_fwd("conv3d", 3, input, kernel, stride, None, padding, "cuboid_im2col", out_dtype, out=out)


def conv3d_backward_input(input, kernel, stride, output_grad, out_dtype=None, out=None):
    """owl_dense_ndarray_generic.ml:3017."""
    return _conv_bwd("conv3d_backward_input", 3, "input", input, kernel, stride, None, output_grad,
                     "cuboid_backward_input_im2col", out_dtype, out=out)


def conv3d_backward_kernel(input, kernel, stride, output_grad, out_dtype=None, out=None):
    return _conv_bwd("conv3d_backward_kernel", 3, "kernel", input, kernel, stride, None, output_grad,
                     "cuboid_backward_kernel_im2col", out_dtype, out=out)


def dilated_conv2d(input, kernel, stride, rate, padding=SAME, out_dtype=None, out=None):
    """owl_dense_ndarray_generic.ml:3797-3880 (output shape uses the dilated extent)."""
    return _conv_fwd("dilated_conv2d", 2, input, kernel, stride, rate, padding, "dilated_spatial_im2col",
                     out_dtype, out=out)


def dilated_conv2d_backward_input(input, kernel, stride, rate, output_grad, out_dtype=None, out=None):
    return _conv_bwd("dilated_conv2d_backward_input", 2, "input", input, kernel, stride, rate, output_grad,
                     "dilated_spatial_backward_input_im2col", out_dtype, out=out)


def dilated_conv2d_backward_kernel(input, kernel, stride, rate, output_grad, out_dtype=None, out=None):
    return _conv_bwd("dilated_conv2d_backward_kernel", 2, "kernel", input, kernel, stride, rate,
                     output_grad, "dilated_spatial_backward_kernel_im2col", out_dtype, out=out)


def dilated_conv3d(input, kernel, stride, rate, padding=SAME, out_dtype=None, out=None):
    return _conv_fwd("dilated_conv3d", 3, input, kernel, stride, rate, padding, "dilated_cuboid_im2col",
                     out_dtype, out=out)


def dilated_conv3d_backward_input(input, kernel, stride, rate, output_grad, out_dtype=None, out=None):
    return _conv_bwd("dilated_conv3d_backward_input", 3, "input", input, kernel, stride, rate, output_grad,
                     "dilated_cuboid_backward_input_im2col", out_dtype, out=out)


def dilated_conv3d_backward_kernel(input, kernel, stride, rate, output_grad, out_dtype=None, out=None):
    return _conv_bwd("dilated_conv3d_backward_kernel", 3, "kernel", input, kernel, stride, rate,
                     output_grad, "dilated_cuboid_backward_kernel_im2col", out_dtype, out=out)


# ----------------------------------------------------------------------------- transpose conv
# Owl runs transpose convolutions on the ordinary kernels with the roles of
# input and output swapped and the kernel buffer [kh; kw; ci; co] read as
# [kh; kw; co; ci] -- a reinterpretation, not a transpose
# (owl_dense_ndarray_generic.ml:5371-5440, 5514-5611, 5713-5806).
def _tconv_geom(x, k, sp_out):
    B, sp_in, Ci = x.shape[0], list(x.shape[1:-1]), x.shape[-1]
    ksp, Co = list(k.shape[:-2]), k.shape[-1]
    return B, sp_in, Ci, ksp, Co


def transpose_conv2d(input, kernel, stride, padding=SAME, out=None):
    return _transpose_fwd("transpose_conv2d", 2, input, kernel, stride, padding, out)


def transpose_conv3d(input, kernel, stride, padding=SAME, out=None):
    return _transpose_fwd("transpose_conv3d", 3, input, kernel, stride, padding, out)


def _transpose_fwd(name, nd, x, k, stride, padding, out=None):
    _same_kind(x, k)
    _check_conv_args(name, nd, x, k, stride)
    x, k = _contig(x), _contig(k)
    B, sp_in, Ci, ksp, Co = _tconv_geom(x, k, None)
    sp_out = [transpose_out_dim(padding, n, kk, s) for n, kk, s in zip(sp_in, ksp, stride)]
    y = _out_buf(out, x, [B, *sp_out, Co], None, name, (x, k))
    # backward_input of a conv whose input is y ([.., Co]) and output is x ([.., Ci])
    pads = _pads_from_shapes(sp_out, ksp, sp_in, stride, [1] * nd)
    g = _Conv(B, sp_out, Co, ksp, Ci, sp_in, stride, [1] * nd, pads)
    if _is_torch(x):
        return _dev_conv("bi", g, k, x, y)
    if nd == 2:
        _host_conv2d("spatial_backward_input", y, k, x, g, None)
    else:
        _host_conv3d("cuboid_backward_input_im2col", y, k, x, g, None)
    return y


def transpose_conv2d_backward_input(input, kernel, stride, output_grad, out=None):
    return _transpose_bwd("transpose_conv2d_backward_input", 2, "input", input, kernel, stride, output_grad,
                          out)


def transpose_conv2d_backward_kernel(input, kernel, stride, output_grad, out=None):
    return _transpose_bwd("transpose_conv2d_backward_kernel", 2, "kernel", input, kernel, stride, output_grad,
                          out)


def transpose_conv3d_backward_input(input, kernel, stride, output_grad, out=None):
    return _transpose_bwd("transpose_conv3d_backward_input", 3, "input", input, kernel, stride, output_grad,
                          out)


def transpose_conv3d_backward_kernel(input, kernel, stride, output_grad, out=None):
    return _transpose_bwd("transpose_conv3d_backward_kernel", 3, "kernel", input, kernel, stride, output_grad,
                          out)


def _transpose_bwd(name, nd, which, x, k, stride, dy, out=None, workspace=None):
    _same_kind(x, k, dy)
    _check_conv_args(name, nd, x, k, stride, None, dy)
    x, k, dy = _contig(x), _contig(k), _contig(dy)
    g = _tbwd_geom(nd, list(x.shape), list(k.shape), stride, list(dy.shape))
    if which == "input":
        res = _out_buf(out, x, list(x.shape), None, name, (x, k, dy))
        if _is_torch(x):
            return _dev_conv("fwd", g, dy, k, res)
        if nd == 2:
            # Owl passes a dummy SAME code here (owl_dense_ndarray_generic.ml:5783)
            _host_conv2d("spatial", dy, k, res, g, 0)
        else:
            _host_conv3d("cuboid_im2col", dy, k, res, g, 0)
    else:
        res = _out_buf(out, k, list(k.shape), None, name, (x, k, dy))
        if _is_torch(x):
            return _dev_conv("bk", g, dy, x, res, workspace=workspace)
        if nd == 2:
            _host_conv2d("spatial_backward_kernel", dy, res, x, g, None)
        else:
            _host_conv3d("cuboid_backward_kernel_im2col", dy, res, x, g, None)
    return res


# ----------------------------------------------------------------------------- pooling
def _pool_check(name, nd, x, kernel, stride):
    _verify(len(x.shape) == nd + 2, f"{name}: input dimension = {len(x.shape)} (should be {nd + 2})")
    _verify(len(kernel) == nd, f"{name}: kernel dimension = {len(kernel)} (should be {nd})")
    _verify(len(stride) == nd, f"{name}: stride dimension = {len(stride)} (should be {nd})")


def _pool_geom(x, kernel, stride, padding):
    B, sp_in, C = x.shape[0], list(x.shape[1:-1]), x.shape[-1]
    sp_out = [conv_out_dim(padding, n, k, s) for n, k, s in zip(sp_in, kernel, stride)]
    if padding == SAME:
        pads = [c_pad(s, o, k, n) for n, k, o, s in zip(sp_in, kernel, sp_out, stride)]
    else:
        pads = [0] * len(sp_in)
    return B, sp_in, C, sp_out, pads


def _pool_desc(B, sp_in, C, kernel, sp_out, stride, pads):
    f = lambda v, d: list(v) + [d] * (3 - len(v))
    return _lib.PoolDesc(B, *f(sp_in, 1), C, *f(kernel, 1), *f(sp_out, 1), *f(stride, 1), *f(pads, 0))


def _pool_fwd(name, nd, is_max, x, kernel, stride, padding, out=None):
    _pool_check(name, nd, x, kernel, stride)
    x = _contig(x)
    B, sp_in, C, sp_out, pads = _pool_geom(x, kernel, stride, padding)
    y = _out_buf(out, x, [B, *sp_out, C], None, name, (x,))
    if _is_torch(x):
        d = _pool_desc(B, sp_in, C, kernel, sp_out, stride, pads)
        _lib.check(_lib.lib().owl_amd_pool_fwd(ctypes.byref(d), _tdt(x), int(is_max), _ptr(x), _ptr(y),
                                               _stream()))
        return y
    pc = 0 if padding == SAME else 1
    op = ("maxpool" if is_max else "avgpool") + ("_spatial" if nd == 2 else "_cuboid")
    if nd == 2:
        # owl_ndarray_pool_impl.h:9-15
        _lib.call(f"owl_amd_{_np_prefix(x)}_ndarray_{op}", _ptr(x), _ptr(y), B, sp_in[0], sp_in[1], C,
                  kernel[0], kernel[1], sp_out[0], sp_out[1], stride[1], stride[0], pc, 1, 1)
    else:
        # owl_ndarray_pool_impl.h:212-219 (strides dpt, row, col)
        _lib.call(f"owl_amd_{_np_prefix(x)}_ndarray_{op}", _ptr(x), _ptr(y), B, *sp_in, C, *kernel,
                  *sp_out, stride[2], stride[1], stride[0], pc)
    return y


def max_pool2d(input, kernel, stride, padding=SAME, out=None):
    """owl_dense_ndarray_generic.ml:6905-6961."""
    return _pool_fwd("max_pool2d", 2, True, input, kernel, stride, padding, out)


def avg_pool2d(input, kernel, stride, padding=SAME, out=None):
    return _pool_fwd("avg_pool2d", 2, False, input, kernel, stride, padding, out)


def max_pool3d(input, kernel, stride, padding=SAME, out=None):
    return _pool_fwd("max_pool3d", 3, True, input, kernel, stride, padding, out)


def avg_pool3d(input, kernel, stride, padding=SAME, out=None):
    return _pool_fwd("avg_pool3d", 3, False, input, kernel, stride, padding, out)


def max_pool2d_argmax(input, kernel, stride, padding=SAME):
    """owl_dense_ndarray_generic.ml:7517-7581: (output, int64 flat argmax)."""
    _pool_check("max_pool2d_argmax", 2, input, kernel, stride)
    x = _contig(input)
    B, (H, W), C = x.shape[0], x.shape[1:3], x.shape[3]
    Ho, Wo = (conv_out_dim(padding, n, k, s) for n, k, s in zip((H, W), kernel, stride))
    top, left, _, _ = calc_conv2d_padding(H, W, kernel[0], kernel[1], Ho, Wo, stride[1], stride[0])
    y = _empty_like_kind(x, [B, Ho, Wo, C])
    if _is_torch(x):
        arg = torch.empty([B, Ho, Wo, C], dtype=torch.int64, device=x.device)
        d = _pool_desc(B, [H, W], C, kernel, [Ho, Wo], stride, [left, top])
        _lib.check(_lib.lib().owl_amd_pool_argmax(ctypes.byref(d), _tdt(x), _ptr(x), _ptr(y), _ptr(arg),
                                                  _stream()))
        return y, arg
    arg = np.empty([B, Ho, Wo, C], dtype=np.int64)
    _lib.call(f"owl_amd_{_np_prefix(x)}_ndarray_maxpool_spatial_arg", _ptr(x), _ptr(y), _ptr(arg), B, H, W,
              C, kernel[0], kernel[1], Ho, Wo, stride[1], stride[0], top, left)
    return y, arg


def _pool_bwd(name, nd, is_max, padding, x, kernel, stride, dy, out=None):
    _pool_check(name, nd, x, kernel, stride)
    _same_kind(x, dy)
    x, dy = _contig(x), _contig(dy)
    B, sp_in, C, sp_out, _ = _pool_geom(x, kernel, stride, padding)
    _verify(list(dy.shape) == [B, *sp_out, C], f"{name}: output' shape {list(dy.shape)} does not match")
    dx = _out_buf(out, x, list(x.shape), None, name, (x, dy))
    if nd == 2:
        top, left, _, _ = calc_conv2d_padding(sp_in[0], sp_in[1], kernel[0], kernel[1], sp_out[0],
                                              sp_out[1], stride[1], stride[0])
        pads = [left, top]
    else:
        pads = [c_pad(s, o, k, n) for n, k, o, s in zip(sp_in, kernel, sp_out, stride)] \
            if padding == SAME else [0, 0, 0]
    if _is_torch(x):
        d = _pool_desc(B, sp_in, C, kernel, sp_out, stride, pads)
        _lib.check(_lib.lib().owl_amd_pool_bwd(ctypes.byref(d), _tdt(x), int(is_max), _ptr(x), _ptr(dy),
                                               _ptr(dx), _stream()))
        return dx
    op = ("maxpool" if is_max else "avgpool") + ("_spatial" if nd == 2 else "_cuboid") + "_backward"
    # Owl's average backward passes the output buffer as the dummy input
    # (owl_ndarray_pool.ml:85-119)
    src = x if is_max else dx
    if nd == 2:
        # owl_ndarray_pool_impl.h:102-108 (pad_rows = top, pad_cols = left)
        _lib.call(f"owl_amd_{_np_prefix(x)}_ndarray_{op}", _ptr(src), _ptr(dy), _ptr(dx), B, sp_in[0],
                  sp_in[1], C, kernel[0], kernel[1], sp_out[0], sp_out[1], stride[1], stride[0],
                  pads[1], pads[0])
    else:
        # owl_ndarray_pool_impl.h:324-331 (strides col, row, dpt; padding code)
        pc = 0 if padding == SAME else 1
        _lib.call(f"owl_amd_{_np_prefix(x)}_ndarray_{op}", _ptr(src), _ptr(dy), _ptr(dx), B, *sp_in, C,
                  *kernel, *sp_out, stride[0], stride[1], stride[2], pc)
    return dx


def max_pool2d_backward(padding, input, kernel, stride, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:7715-7775."""
    return _pool_bwd("max_pool2d_backward", 2, True, padding, input, kernel, stride, output_grad, out)


def avg_pool2d_backward(padding, input, kernel, stride, output_grad, out=None):
    return _pool_bwd("avg_pool2d_backward", 2, False, padding, input, kernel, stride, output_grad, out)


def max_pool3d_backward(padding, input, kernel, stride, output_grad, out=None):
    return _pool_bwd("max_pool3d_backward", 3, True, padding, input, kernel, stride, output_grad, out)


def avg_pool3d_backward(padding, input, kernel, stride, output_grad, out=None):
    return _pool_bwd("avg_pool3d_backward", 3, False, padding, input, kernel, stride, output_grad, out)


# ----------------------------------------------------------------------------- fused conv layer epilogue
_ACT = {None: 0, "none": 0, "relu": 1}


def _act_code(act):
    _verify(act in _ACT, f"unknown activation {act!r} (none or relu)")
    return _ACT[act]


def _conv_bias_act(name, nd, x, k, bias, stride, rate, padding, act, keep_pre, out_dtype, out=None, pre_out=None):
    _verify(_is_torch(x) and _is_torch(k), f"{name}: the fused layer ops take device tensors")
    _check_conv_args(name, nd, x, k, stride, rate)
    x, k = _contig(x), _contig(k)
    g = conv_problem(list(x.shape), list(k.shape), stride, rate, padding)
    y = _out_buf(out, x, [g.B, *g.sp_out, g.Co], out_dtype, name, (x, k))
    pre = _out_buf(pre_out, x, list(y.shape), y.dtype, name, (x, k, y)) if keep_pre else None
    acc_dt = torch.float64 if y.dtype == torch.float64 else torch.float32
    if bias is not None:
        _verify(_is_torch(bias) and tuple(bias.shape) == (g.Co,),
                f"{name}: bias must be a device tensor of shape [{g.Co}]")
        _verify(bias.dtype == acc_dt, f"{name}: bias must be {acc_dt} for {y.dtype} output")
        bias = _contig(bias)
    d = g.desc()
    _lib.check(_lib.lib().owl_amd_conv_fwd_bias_act(
        ctypes.byref(d), _tdt(x), _tdt(y), _ptr(x), _ptr(k), _ptr(bias) if bias is not None else None,
        _act_code(act), _ptr(y), _ptr(pre) if pre is not None else None, _stream()))
    return (y, pre) if keep_pre else y


def conv2d_bias_act(input, kernel, bias, stride, padding=SAME, act="relu", keep_pre=False, out_dtype=None,
                    out=None, pre_out=None):
    """Owl's 2-D conv layer forward, ``act (conv2d ~padding x w stride + b)``
    (owl_neural_neuron.ml:1253 Conv2D.run, then the activation), with the bias
    add and the activation in the convolution's epilogue.  relu is
    ``fmax(v, 0)`` as owl_ndarray_maths_stub.c:3559-3569.  Device tensors
    only; bias is [Co] in float32 (float64 for float64 output) or None.
    Returns y, or (y, pre) with pre = conv + bias when keep_pre; ``out`` and
    ``pre_out`` are optional arrays to write them into."""
    return _conv_bias_act("conv2d_bias_act", 2, input, kernel, bias, stride, None, padding, act, keep_pre,
                          out_dtype, out, pre_out)


def conv2d_bias_act_maxpool(input, kernel, bias, stride, padding=SAME, act="relu", keep_y=True, keep_pre=False,
                            out_dtype=None, out=None):
    """Owl's conv layer followed by a 2x2 stride-2 VALID max-pool layer (a VGG
    block's end: owl_neural_neuron.ml:1253 Conv2D.run and the activation, then
    :2024 MaxPool2D.run = ``max_pool2d ~padding:VALID x [|2;2|] [|2;2|]``).
    Returns (y, pooled), or (y, pooled, pre) when keep_pre; y is None when
    keep_y is False (inference: only the pooled layer leaves the chip).
    pooled equals ``max_pool2d(y, [2, 2], [2, 2], VALID)`` bit for bit; for
    BF16 / FP16 stride-1 layers it is taken from the convolution's output
    tile before that is stored.  ``out`` is an optional array for pooled."""
    name = "conv2d_bias_act_maxpool"
    _verify(_is_torch(input) and _is_torch(kernel), f"{name}: the fused layer ops take device tensors")
    _check_conv_args(name, 2, input, kernel, stride, None)
    x, k = _contig(input), _contig(kernel)
    g = conv_problem(list(x.shape), list(k.shape), stride, None, padding)
    ys = [g.B, *g.sp_out, g.Co]
    y = _out_buf(None, x, ys, out_dtype, name, (x, k)) if keep_y else None
    odt = y.dtype if y is not None else (out_dtype or x.dtype)
    pre = _out_buf(None, x, ys, odt, name, (x, k)) if keep_pre else None
    pooled = _out_buf(out, x, [g.B, g.sp_out[0] // 2, g.sp_out[1] // 2, g.Co], odt, name, (x, k))
    acc_dt = torch.float64 if odt == torch.float64 else torch.float32
    if bias is not None:
        _verify(_is_torch(bias) and tuple(bias.shape) == (g.Co,),
                f"{name}: bias must be a device tensor of shape [{g.Co}]")
        _verify(bias.dtype == acc_dt, f"{name}: bias must be {acc_dt} for {odt} output")
        bias = _contig(bias)
    d = g.desc()
    _lib.check(_lib.lib().owl_amd_conv_fwd_bias_act_pool(
        ctypes.byref(d), _tdt(x), _tdt(pooled), _ptr(x), _ptr(k), _ptr(bias) if bias is not None else None,
        _act_code(act), _ptr(y) if y is not None else None, _ptr(pre) if pre is not None else None,
        _ptr(pooled), _stream()))
    return (y, pooled, pre) if keep_pre else (y, pooled)


def conv3d_bias_act(input, kernel, bias, stride, padding=SAME, act="relu", keep_pre=False, out_dtype=None,
                    out=None, pre_out=None):
    """The 3-D form of ``conv2d_bias_act`` (Owl's Conv3D layer)."""
    return _conv_bias_act("conv3d_bias_act", 3, input, kernel, bias, stride, None, padding, act, keep_pre,
                          out_dtype, out, pre_out)


def bias_act_backward(pre, output_grad, act="relu", bias_grad=True):
    """Adjoint of ``y = act(pre)`` with ``pre = conv + b``: returns (dz, db).
    dz = dy * (signum(pre) + 1) / 2 for relu -- Owl's relu adjoint
    (owl_algodiff_ops.ml:774), 1 above zero, 1/2 at zero, 0 below -- or dy
    for no activation; db sums dz over every axis but the last (float32, or
    float64 for float64 tensors), in a fixed order.  Device tensors only."""
    _verify(_is_torch(output_grad), "bias_act_backward takes device tensors")
    code = _act_code(act)
    dy = _contig(output_grad)
    if code:
        _verify(pre is not None and _is_torch(pre) and pre.shape == dy.shape and pre.dtype == dy.dtype,
                "bias_act_backward: pre must match the output gradient")
        pre = _contig(pre)
    C = dy.shape[-1]
    npix = dy.numel() // C if C else 0
    dz = torch.empty_like(dy)
    db = None
    if bias_grad:
        db = torch.empty(C, dtype=torch.float64 if dy.dtype == torch.float64 else torch.float32, device=dy.device)
    _lib.check(_lib.lib().owl_amd_bias_act_bwd(
        _tdt(dy), code, npix, C, _ptr(pre) if code else None, _ptr(dy), _ptr(dz),
        _ptr(db) if db is not None else None, _stream()))
    return dz, db


def maxpool_bias_act_backward(y, pre, pooled_grad, bias_grad=True):
    """Adjoint of ``pooled = max_pool2d(y, [2, 2], [2, 2], VALID)`` with
    ``y = relu(pre)`` (the backward of ``conv2d_bias_act_maxpool``): returns
    (dz, db).  dz has the bits of ``max_pool2d_backward`` into dy followed by
    ``bias_act_backward(pre, dy)``; for even H and W one pass forms it without
    the intermediate dy (db then summed in another fixed order), otherwise
    those two calls run.  y and pre are [B, H, W, C], pooled_grad
    [B, H // 2, W // 2, C]; device tensors only."""
    _verify(_is_torch(y) and _is_torch(pre) and _is_torch(pooled_grad),
            "maxpool_bias_act_backward takes device tensors")
    _verify(y.dim() == 4 and pre.shape == y.shape and pre.dtype == y.dtype and pooled_grad.dtype == y.dtype,
            "maxpool_bias_act_backward: y and pre must be [B, H, W, C] of one dtype")
    B, H, W, C = y.shape
    _verify(list(pooled_grad.shape) == [B, H // 2, W // 2, C],
            f"maxpool_bias_act_backward: pooled_grad shape {list(pooled_grad.shape)} should be {[B, H // 2, W // 2, C]}")
    if H % 2 or W % 2:
        return bias_act_backward(pre, max_pool2d_backward(VALID, y, [2, 2], [2, 2], pooled_grad), "relu", bias_grad)
    y, pre, dp = _contig(y), _contig(pre), _contig(pooled_grad)
    dz = torch.empty_like(y)
    db = None
    if bias_grad:
        db = torch.empty(C, dtype=torch.float64 if y.dtype == torch.float64 else torch.float32, device=y.device)
    _lib.check(_lib.lib().owl_amd_maxpool_bias_act_bwd(
        _tdt(y), _act_code("relu"), B, H, W, C, _ptr(y), _ptr(pre), _ptr(dp), _ptr(dz),
        _ptr(db) if db is not None else None, _stream()))
    return dz, db


# ----------------------------------------------------------------------------- upsampling
def _check_upsampling(name, x, size):
    _verify(len(x.shape) == 4 and len(size) == 2,
            f"{name}: input dimension = {len(x.shape)} (should be 4); "
            f"size dimension = {len(size)} (should be 2).")


def upsampling2d(input, size, out=None):
    """owl_dense_ndarray_generic.ml:8223-8234: ``repeat input [|1; s0; s1; 1|]``
    (nearest neighbour, s0 along H = Owl's cols, s1 along W = rows)."""
    _check_upsampling("upsampling2d", input, size)
    x = _contig(input)
    B, H, W, C = x.shape
    s0, s1 = int(size[0]), int(size[1])
    y = _out_buf(out, x, [B, H * s0, W * s1, C], None, "upsampling2d", (x,))
    if _is_torch(x):
        _lib.check(_lib.lib().owl_amd_upsampling2d_fwd(_tdt(x), B, H, W, C, H * s0, W * s1, s0, s1, _ptr(x),
                                                        _ptr(y), _stream()))
        return y
    _lib.call(f"owl_amd_{_np_prefix(x)}_ndarray_upsampling_spatial", _ptr(x), _ptr(y), B, H, W, C, H * s0,
              W * s1, s0, s1)
    return y


def upsampling2d_backward(input, size, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:8249-8290 and the stub
    owl_ndarray_upsampling_impl.h:9-60: the input's gradient, each cell the
    sum of the output gradient over its s0 x s1 block (Owl's order)."""
    _check_upsampling("upsampling2d_backward", input, size)
    _same_kind(input, output_grad)
    B, H, W, C = input.shape
    s0, s1 = int(size[0]), int(size[1])
    dy = _contig(output_grad)
    _verify(dy.shape[1] == H * s0 and dy.shape[2] == W * s1,
            f"upsampling2d_backward: output shape is [{','.join(map(str, dy.shape))}]; scaled output cols is "
            f"{H * s0}, should be equal to the 2nd dimension of output shape; scaled output rows is {W * s1}, "
            f"should be equal to the 3rd dimension of output shape.")
    dx = _out_buf(out, dy, [B, H, W, C], None, "upsampling2d_backward", (dy,))
    if _is_torch(dy):
        _lib.check(_lib.lib().owl_amd_upsampling2d_bwd(_tdt(dy), B, H, W, C, H * s0, W * s1, s0, s1, _ptr(dy),
                                                        _ptr(dx), _stream()))
        return dx
    _lib.call(f"owl_amd_{_np_prefix(dy)}_ndarray_upsampling_spatial_backward", _ptr(dx), _ptr(dy), B, H, W, C,
              H * s0, W * s1, s0, s1)
    return dx


# ----------------------------------------------------------------------------- 1-D convolution and pooling
# Owl runs every 1-D op as the 2-D one on [B; 1; W; C] with kernel [1; kw; Ci; Co]
# and stride [1; s] (owl_dense_ndarray_generic.ml:3402-3787 conv, 4967-5363
# dilated, 6483-6890 transpose, 7027-7230 / 7840-8224 pooling); so does this.
def _rs(a, shape):
    return a.reshape(shape)


def _check_1d(name, x, k=None, stride=None, dy=None, kernel_len=None):
    _verify(len(x.shape) == 3, f"{name}: input dimension = {len(x.shape)} (should be 3)")
    if k is not None:
        _verify(len(k.shape) == 3, f"{name}: kernel dimension = {len(k.shape)} (should be 3)")
        _verify(x.shape[2] == k.shape[1],
                f"{name}: the 3rd dimension of input shape should be equal to the 2nd dimension of kernel shape")
    if kernel_len is not None:
        _verify(kernel_len == 1, f"{name}: kernel dimension = {kernel_len} (should be 1)")
    if stride is not None:
        _verify(len(stride) == 1, f"{name}: stride dimension = {len(stride)} (should be 1)")
    if dy is not None:
        _verify(len(dy.shape) == 3, f"{name}: output' dimension = {len(dy.shape)} (should be 3)")


def _x2(x):
    return _rs(x, [x.shape[0], 1, x.shape[1], x.shape[2]])


def _k2(k):
    return _rs(k, [1, *k.shape])


def _y1(y):
    return _rs(y, [y.shape[0], y.shape[2], y.shape[3]])


def _o2(out, name, kernel=False):
    """A 1-D op's ``out`` as the 2-D view its 2-D op writes (a reshape that
    must not copy, so the array has to be contiguous)."""
    if out is None:
        return None
    _verify(len(out.shape) == 3, f"{name}: out dimension = {len(out.shape)} (should be 3)")
    _verify(out.is_contiguous() if _is_torch(out) else out.flags.c_contiguous, f"{name}: out must be contiguous")
    return _k2(out) if kernel else _x2(out)


def conv1d(input, kernel, stride, padding=SAME, out=None):
    """owl_dense_ndarray_generic.ml:3402-3445."""
    _check_1d("conv1d", input, kernel, stride)
    return _y1(conv2d(_x2(input), _k2(kernel), [1, stride[0]], padding, out=_o2(out, "conv1d")))


def conv1d_backward_input(input, kernel, stride, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:3489-3561."""
    _check_1d("conv1d_backward_input", input, kernel, stride, output_grad)
    return _rs(conv2d_backward_input(_x2(input), _k2(kernel), [1, stride[0]], _x2(output_grad),
                                     out=_o2(out, "conv1d_backward_input")), input.shape)


def conv1d_backward_kernel(input, kernel, stride, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:3639-3713."""
    _check_1d("conv1d_backward_kernel", input, kernel, stride, output_grad)
    return _rs(conv2d_backward_kernel(_x2(input), _k2(kernel), [1, stride[0]], _x2(output_grad),
                                      out=_o2(out, "conv1d_backward_kernel", kernel=True)), kernel.shape)


def dilated_conv1d(input, kernel, stride, rate, padding=SAME, out=None):
    """owl_dense_ndarray_generic.ml:4967-5010.  Owl hands ``rate`` to the 2-D
    function unchanged, so it must have two entries (rows, cols) as there."""
    _check_1d("dilated_conv1d", input, kernel, stride)
    return _y1(dilated_conv2d(_x2(input), _k2(kernel), [1, stride[0]], rate, padding,
                              out=_o2(out, "dilated_conv1d")))


def dilated_conv1d_backward_input(input, kernel, stride, rate, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:5054-5130 (``rate`` as in dilated_conv1d)."""
    _check_1d("dilated_conv1d_backward_input", input, kernel, stride, output_grad)
    return _rs(dilated_conv2d_backward_input(_x2(input), _k2(kernel), [1, stride[0]], rate, _x2(output_grad),
                                             out=_o2(out, "dilated_conv1d_backward_input")), input.shape)


def dilated_conv1d_backward_kernel(input, kernel, stride, rate, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:5210-5286 (``rate`` as in dilated_conv1d)."""
    _check_1d("dilated_conv1d_backward_kernel", input, kernel, stride, output_grad)
    return _rs(dilated_conv2d_backward_kernel(_x2(input), _k2(kernel), [1, stride[0]], rate, _x2(output_grad),
                                              out=_o2(out, "dilated_conv1d_backward_kernel", kernel=True)),
               kernel.shape)


def transpose_conv1d(input, kernel, stride, padding=SAME, out=None):
    """owl_dense_ndarray_generic.ml:6483-6526."""
    _check_1d("transpose_conv1d", input, kernel, stride)
    return _y1(transpose_conv2d(_x2(input), _k2(kernel), [1, stride[0]], padding, out=_o2(out, "transpose_conv1d")))


def transpose_conv1d_backward_input(input, kernel, stride, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:6570-6646."""
    _check_1d("transpose_conv1d_backward_input", input, kernel, stride, output_grad)
    return _rs(transpose_conv2d_backward_input(_x2(input), _k2(kernel), [1, stride[0]], _x2(output_grad),
                                               out=_o2(out, "transpose_conv1d_backward_input")), input.shape)


def transpose_conv1d_backward_kernel(input, kernel, stride, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:6732-6814."""
    _check_1d("transpose_conv1d_backward_kernel", input, kernel, stride, output_grad)
    return _rs(transpose_conv2d_backward_kernel(_x2(input), _k2(kernel), [1, stride[0]], _x2(output_grad),
                                                out=_o2(out, "transpose_conv1d_backward_kernel", kernel=True)),
               kernel.shape)


def max_pool1d(input, kernel, stride, padding=SAME, out=None):
    """owl_dense_ndarray_generic.ml:7027-7053."""
    _check_1d("max_pool1d", input, stride=stride, kernel_len=len(kernel))
    return _y1(max_pool2d(_x2(input), [1, kernel[0]], [1, stride[0]], padding, out=_o2(out, "max_pool1d")))


def avg_pool1d(input, kernel, stride, padding=SAME, out=None):
    """owl_dense_ndarray_generic.ml:7197-7223."""
    _check_1d("avg_pool1d", input, stride=stride, kernel_len=len(kernel))
    return _y1(avg_pool2d(_x2(input), [1, kernel[0]], [1, stride[0]], padding, out=_o2(out, "avg_pool1d")))


def max_pool1d_backward(padding, input, kernel, stride, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:7840-7871."""
    _check_1d("max_pool1d_backward", input, stride=stride, kernel_len=len(kernel))
    return _rs(max_pool2d_backward(padding, _x2(input), [1, kernel[0]], [1, stride[0]], _x2(output_grad),
                                   out=_o2(out, "max_pool1d_backward")), input.shape)


def avg_pool1d_backward(padding, input, kernel, stride, output_grad, out=None):
    """owl_dense_ndarray_generic.ml:8158-8189."""
    _check_1d("avg_pool1d_backward", input, stride=stride, kernel_len=len(kernel))
    return _rs(avg_pool2d_backward(padding, _x2(input), [1, kernel[0]], [1, stride[0]], _x2(output_grad),
                                   out=_o2(out, "avg_pool1d_backward")), input.shape)


# ----------------------------------------------------------------------------- in-place variants
# Owl's ``*_ ~out`` wrappers compute into a given array (the computation
# graph's evaluator calls these, owl_computation_cpu_eval.ml:180-281).  Here
# ``out`` is the first argument, then the function's own arguments; the array
# is checked (kind, shape, dtype, contiguity, no overlap with an input) and
# the call returns None.
def _inplace(f, owl_line):
    def g(out, *args, **kw):
        f(*args, out=out, **kw)
    g.__name__ = g.__qualname__ = f.__name__ + "_"
    g.__doc__ = (f"``{f.__name__}`` written into ``out`` (``{f.__name__}_ ~out``, "
                 f"owl_dense_ndarray_generic.ml:{owl_line}).")
    return g


conv1d_ = _inplace(conv1d, 3447)
conv1d_backward_input_ = _inplace(conv1d_backward_input, 3563)
conv1d_backward_kernel_ = _inplace(conv1d_backward_kernel, 3715)
conv2d_ = _inplace(conv2d, 2402)
conv2d_backward_input_ = _inplace(conv2d_backward_input, 2569)
conv2d_backward_kernel_ = _inplace(conv2d_backward_kernel, 2752)
conv3d_ = _inplace(conv3d, 2934)
conv3d_backward_input_ = _inplace(conv3d_backward_input, 3111)
conv3d_backward_kernel_ = _inplace(conv3d_backward_kernel, 3302)
dilated_conv1d_ = _inplace(dilated_conv1d, 5012)
dilated_conv1d_backward_input_ = _inplace(dilated_conv1d_backward_input, 5132)
dilated_conv1d_backward_kernel_ = _inplace(dilated_conv1d_backward_kernel, 5288)
dilated_conv2d_ = _inplace(dilated_conv2d, 3879)
dilated_conv2d_backward_input_ = _inplace(dilated_conv2d_backward_input, 4056)
dilated_conv2d_backward_kernel_ = _inplace(dilated_conv2d_backward_kernel, 4247)
dilated_conv3d_ = _inplace(dilated_conv3d, 4445)
dilated_conv3d_backward_input_ = _inplace(dilated_conv3d_backward_input, 4645)
dilated_conv3d_backward_kernel_ = _inplace(dilated_conv3d_backward_kernel, 4856)
transpose_conv1d_ = _inplace(transpose_conv1d, 6528)
transpose_conv1d_backward_input_ = _inplace(transpose_conv1d_backward_input, 6648)
transpose_conv1d_backward_kernel_ = _inplace(transpose_conv1d_backward_kernel, 6816)
transpose_conv2d_ = _inplace(transpose_conv2d, 5443)
transpose_conv2d_backward_input_ = _inplace(transpose_conv2d_backward_input, 5809)
transpose_conv2d_backward_kernel_ = _inplace(transpose_conv2d_backward_kernel, 5614)
transpose_conv3d_ = _inplace(transpose_conv3d, 5995)
transpose_conv3d_backward_input_ = _inplace(transpose_conv3d_backward_input, 6172)
transpose_conv3d_backward_kernel_ = _inplace(transpose_conv3d_backward_kernel, 6381)
max_pool1d_ = _inplace(max_pool1d, 7055)
max_pool2d_ = _inplace(max_pool2d, 6964)
max_pool3d_ = _inplace(max_pool3d, 7322)
avg_pool1d_ = _inplace(avg_pool1d, 7225)
avg_pool2d_ = _inplace(avg_pool2d, 7139)
avg_pool3d_ = _inplace(avg_pool3d, 7453)
max_pool1d_backward_ = _inplace(max_pool1d_backward, 7873)
max_pool2d_backward_ = _inplace(max_pool2d_backward, 7778)
max_pool3d_backward_ = _inplace(max_pool3d_backward, 7650)
avg_pool1d_backward_ = _inplace(avg_pool1d_backward, 8191)
avg_pool2d_backward_ = _inplace(avg_pool2d_backward, 8097)
avg_pool3d_backward_ = _inplace(avg_pool3d_backward, 7971)
upsampling2d_ = _inplace(upsampling2d, 8236)
upsampling2d_backward_ = _inplace(upsampling2d_backward, 8305)

__all__ += [n for n in list(globals()) if n.endswith("_") and not n.startswith("_") and callable(globals()[n])]
