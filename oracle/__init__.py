"""Python access to the CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package; the product path never does.  The library
``oracle/liboracle.so`` is built from ``oracle/owl_oracle.c`` by
``oracle/Makefile`` (``__graft_entry__.build()`` runs it).
"""
import ctypes
import glob
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class ConvGeom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "B", "H", "W", "D", "Ci", "kh", "kw", "kd", "Co", "Ho", "Wo", "Do",
        "sh", "sw", "sd", "rh", "rw", "rd", "ph", "pw", "pd")]


class PoolGeom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "B", "H", "W", "D", "C", "kh", "kw", "kd", "Ho", "Wo", "Do",
        "sh", "sw", "sd", "ph", "pw", "pd")]


SGEMM_T = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                           ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                           ctypes.c_void_p, ctypes.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        for n in ("oracle_conv_fwd_f64", "oracle_conv_bwd_input_f64", "oracle_conv_bwd_kernel_f64"):
            getattr(L, n).argtypes = [ctypes.POINTER(ConvGeom), vp, vp, vp]
            getattr(L, n).restype = None
        for n in ("oracle_conv_fwd_owl_f32", "oracle_conv_bwd_input_owl_f32",
                  "oracle_conv_bwd_kernel_owl_f32"):
            getattr(L, n).argtypes = [ctypes.POINTER(ConvGeom), vp, vp, vp, vp]
            getattr(L, n).restype = ctypes.c_int
            getattr(L, n + "_cache").argtypes = [ctypes.POINTER(ConvGeom), vp, vp, vp, vp, vp]
            getattr(L, n + "_cache").restype = ctypes.c_int
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_owl_block_sizes.argtypes = [ip, ip, ip] + [ctypes.c_int] * 4
        L.oracle_owl_block_sizes.restype = None
        L.oracle_host_cache_sizes.argtypes = [ip, ip, ip]
        L.oracle_host_cache_sizes.restype = None
        for s in ("f32", "f64"):
            getattr(L, "oracle_pool_fwd_" + s).argtypes = [ctypes.POINTER(PoolGeom), ctypes.c_int, vp, vp]
            getattr(L, "oracle_pool_argmax_" + s).argtypes = [ctypes.POINTER(PoolGeom), vp, vp, vp]
            getattr(L, "oracle_pool_bwd_" + s).argtypes = [ctypes.POINTER(PoolGeom), ctypes.c_int, vp, vp, vp]
            for d in ("fwd", "bwd"):
                getattr(L, f"oracle_upsampling_{d}_" + s).argtypes = [vp, vp] + [ctypes.c_int64] * 8
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- geometry
def conv_geom(x_shape, k_shape, out_sp, stride, rate, pads):
    """x_shape [B,H,W,(D,)Ci], k_shape [kh,kw,(kd,)Ci,Co], out_sp (Ho,Wo[,Do])."""
    if len(x_shape) == 4:
        B, H, W, Ci = x_shape
        D = 1
        kh, kw, _, Co = k_shape
        kd = 1
        Ho, Wo = out_sp
        Do = 1
        stride, rate, pads = tuple(stride) + (1,), tuple(rate) + (1,), tuple(pads) + (0,)
    else:
        B, H, W, D, Ci = x_shape
        kh, kw, kd, _, Co = k_shape
        Ho, Wo, Do = out_sp
    return ConvGeom(B, H, W, D, Ci, kh, kw, kd, Co, Ho, Wo, Do, *stride, *rate, *pads)


def pool_geom(x_shape, win, out_sp, stride, pads):
    if len(x_shape) == 4:
        B, H, W, C = x_shape
        D = 1
        win, out_sp = tuple(win) + (1,), tuple(out_sp) + (1,)
        stride, pads = tuple(stride) + (1,), tuple(pads) + (0,)
    else:
        B, H, W, D, C = x_shape
    return PoolGeom(B, H, W, D, C, *win, *out_sp, *stride, *pads)


# ---------------------------------------------------------------- conv
def conv_fwd(g, x, w):
    x = np.ascontiguousarray(x, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    y = np.zeros(g.B * g.Ho * g.Wo * g.Do * g.Co, np.float64)
    lib().oracle_conv_fwd_f64(ctypes.byref(g), _p(x), _p(w), _p(y))
    return y


def conv_bwd_input(g, w, dy):
    w = np.ascontiguousarray(w, np.float64)
    dy = np.ascontiguousarray(dy, np.float64)
    dx = np.zeros(g.B * g.H * g.W * g.D * g.Ci, np.float64)
    lib().oracle_conv_bwd_input_f64(ctypes.byref(g), _p(w), _p(dy), _p(dx))
    return dx


def conv_bwd_kernel(g, x, dy):
    x = np.ascontiguousarray(x, np.float64)
    dy = np.ascontiguousarray(dy, np.float64)
    dw = np.zeros(g.kh * g.kw * g.kd * g.Ci * g.Co, np.float64)
    lib().oracle_conv_bwd_kernel_f64(ctypes.byref(g), _p(x), _p(dy), _p(dw))
    return dw


# ---------------------------------------------------------------- Owl's CPU algorithm
_SGEMM = None


def openblas_sgemm():
    """cblas_sgemm of the OpenBLAS that ships with scipy / numpy in this image."""
    global _SGEMM
    if _SGEMM is None:
        cands = []
        for pat in ("scipy.libs/libscipy_openblas-*.so", "numpy.libs/libscipy_openblas64_*.so"):
            for d in __import__("site").getsitepackages() + ["/usr/local/lib/python3.10/dist-packages"]:
                cands += glob.glob(os.path.join(d, pat))
        for c in cands:
            L = ctypes.CDLL(c)
            for sym in ("scipy_cblas_sgemm", "cblas_sgemm", "scipy_cblas_sgemm64_"):
                f = getattr(L, sym, None)
                if f is not None and "64_" not in sym:
                    _SGEMM = (L, ctypes.cast(f, ctypes.c_void_p), c)
                    break
            if _SGEMM:
                break
        if _SGEMM is None:
            raise RuntimeError("no LP64 OpenBLAS cblas_sgemm found")
    return _SGEMM


def _cache_arg(cache):
    return None if cache is None else (ctypes.c_int * 3)(*cache)


def conv_fwd_owl_f32(g, x, w, cache=None, branch=None):
    """Owl's FP32 forward (im2col + SGEMM, or the GEBP blocks for large 2-D
    layers).  cache: (l1, l2, l3) bytes for the block sizes, else the host's
    CPUID values.  branch: a list that receives "gebp" or "im2col"."""
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    y = np.zeros(g.B * g.Ho * g.Wo * g.Do * g.Co, np.float32)
    rc = lib().oracle_conv_fwd_owl_f32_cache(ctypes.byref(g), _p(x), _p(w), _p(y), openblas_sgemm()[1],
                                             _cache_arg(cache))
    assert rc >= 0
    if branch is not None:
        branch.append("gebp" if rc == 1 else "im2col")
    return y


def conv_bwd_input_owl_f32(g, w, dy, cache=None, branch=None):
    w = np.ascontiguousarray(w, np.float32)
    dy = np.ascontiguousarray(dy, np.float32)
    dx = np.zeros(g.B * g.H * g.W * g.D * g.Ci, np.float32)
    rc = lib().oracle_conv_bwd_input_owl_f32_cache(ctypes.byref(g), _p(w), _p(dy), _p(dx), openblas_sgemm()[1],
                                                   _cache_arg(cache))
    assert rc >= 0
    if branch is not None:
        branch.append("gebp" if rc == 1 else "im2col")
    return dx


def conv_bwd_kernel_owl_f32(g, x, dy, cache=None, branch=None):
    x = np.ascontiguousarray(x, np.float32)
    dy = np.ascontiguousarray(dy, np.float32)
    dw = np.zeros(g.kh * g.kw * g.kd * g.Ci * g.Co, np.float32)
    rc = lib().oracle_conv_bwd_kernel_owl_f32_cache(ctypes.byref(g), _p(x), _p(dy), _p(dw), openblas_sgemm()[1],
                                                    _cache_arg(cache))
    assert rc >= 0
    if branch is not None:
        branch.append("gebp" if rc == 1 else "im2col")
    return dw


def owl_block_sizes(k, m, n, typesize, l1, l2, l3):
    """Owl's compute_block_sizes for (k, m, n) at the given cache sizes."""
    a, b, c = ctypes.c_int(k), ctypes.c_int(m), ctypes.c_int(n)
    lib().oracle_owl_block_sizes(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), typesize, l1, l2, l3)
    return a.value, b.value, c.value


def host_cache_sizes():
    """(l1, l2, l3) bytes as Owl's query_cache_sizes reads them on this host."""
    v = [ctypes.c_int() for _ in range(3)]
    lib().oracle_host_cache_sizes(*[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


# ---------------------------------------------------------------- pooling
def _suf(dt):
    return "f32" if np.dtype(dt) == np.float32 else "f64"


def pool_fwd(g, x, is_max):
    y = np.zeros(g.B * g.Ho * g.Wo * g.Do * g.C, x.dtype)
    x = np.ascontiguousarray(x)
    getattr(lib(), "oracle_pool_fwd_" + _suf(x.dtype))(ctypes.byref(g), int(is_max), _p(x), _p(y))
    return y


def pool_argmax(g, x):
    x = np.ascontiguousarray(x)
    y = np.zeros(g.B * g.Ho * g.Wo * g.Do * g.C, x.dtype)
    a = np.zeros(g.B * g.Ho * g.Wo * g.Do * g.C, np.int64)
    getattr(lib(), "oracle_pool_argmax_" + _suf(x.dtype))(ctypes.byref(g), _p(x), _p(y), _p(a))
    return y, a


def pool_bwd(g, x, dy, is_max):
    x = np.ascontiguousarray(x)
    dy = np.ascontiguousarray(dy, x.dtype)
    dx = np.zeros(g.B * g.H * g.W * g.D * g.C, x.dtype)
    getattr(lib(), "oracle_pool_bwd_" + _suf(x.dtype))(ctypes.byref(g), int(is_max), _p(x), _p(dy), _p(dx))
    return dx


def upsampling_fwd(x, size, out_cols=None, out_rows=None):
    """y = repeat x [|1; size0; size1; 1|] (Owl's upsampling2d), optionally to an
    explicit output extent (clamped to the last input row / column)."""
    x = np.ascontiguousarray(x)
    B, H, W, C = x.shape
    Ho = H * size[0] if out_cols is None else out_cols
    Wo = W * size[1] if out_rows is None else out_rows
    y = np.empty([B, Ho, Wo, C], x.dtype)
    getattr(lib(), "oracle_upsampling_fwd_" + _suf(x.dtype))(_p(x), _p(y), B, H, W, C, Ho, Wo, size[0], size[1])
    return y


def upsampling_bwd(dy, in_shape, size):
    """Owl's upsampling2d_backward stub restated (scatter in Owl's loop order)."""
    dy = np.ascontiguousarray(dy)
    B, H, W, C = in_shape
    dx = np.empty([B, H, W, C], dy.dtype)
    getattr(lib(), "oracle_upsampling_bwd_" + _suf(dy.dtype))(_p(dx), _p(dy), B, H, W, C, dy.shape[1], dy.shape[2],
                                                            size[0], size[1])
    return dx
