"""ctypes binding of ``libowl_amd.so`` (the C ABI declared in include/owl_amd.h).

The library is required: importing a function that needs it raises
``OwlAmdError`` when ``owl_amd/libowl_amd.so`` has not been built.  There is
no CPU fallback anywhere in the package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libowl_amd.so")

OK, ERR_ARG, ERR_HIP, ERR_UNSUPPORTED, ERR_NOMEM = 0, 1, 2, 3, 4
F32, F64, BF16, F16, FP8E4M3, C32, C64 = 0, 1, 2, 3, 4, 5, 6
KT_FWD, KT_BWD_INPUT, KT_BWD_KERNEL, KT_AUX = 0, 1, 2, 3


class OwlAmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"owl_amd error {code}: {msg}")
        self.code = code


class ConvDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "B", "H", "W", "D", "Ci", "kh", "kw", "kd", "Co", "Ho", "Wo", "Do",
        "sh", "sw", "sd", "rh", "rw", "rd", "ph", "pw", "pd")]


class PoolDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "B", "H", "W", "D", "C", "kh", "kw", "kd", "Ho", "Wo", "Do",
        "sh", "sw", "sd", "ph", "pw", "pd")]


_lib = None

# Argument kinds of the drop-in entry points: 'p' data pointer, 'i' int64.
CONV2D_FWD = "ppp" + "i" * 14
CONV2D_BWD = "ppp" + "i" * 13
CONV3D_FWD = "ppp" + "i" * 16
CONV3D_BWD = "ppp" + "i" * 15
DCONV3D_FWD = "ppp" + "i" * 19
DCONV3D_BWD = "ppp" + "i" * 18
POOL2D_FWD = "pp" + "i" * 13
POOL2D_BWD = "ppp" + "i" * 12
POOL2D_ARG = "ppp" + "i" * 12
POOL3D_FWD = "pp" + "i" * 15
POOL3D_BWD = "ppp" + "i" * 15
UPSAMPLE = "pp" + "i" * 8

DROPINS = {}
for _n in ("float32", "float64", "complex32", "complex64"):
    for _v, _k in (("spatial", CONV2D_FWD), ("spatial_im2col", CONV2D_FWD), ("spatial_mec", CONV2D_FWD),
                   ("spatial_naive", CONV2D_FWD), ("dilated_spatial_im2col", CONV2D_FWD),
                   ("spatial_backward_input", CONV2D_BWD), ("spatial_backward_kernel", CONV2D_BWD),
                   ("spatial_backward_input_im2col", CONV2D_BWD), ("spatial_backward_kernel_im2col", CONV2D_BWD),
                   ("spatial_backward_input_mec", CONV2D_BWD), ("spatial_backward_kernel_mec", CONV2D_BWD),
                   ("spatial_backward_input_naive", CONV2D_BWD), ("spatial_backward_kernel_naive", CONV2D_BWD),
                   ("dilated_spatial_backward_input_im2col", CONV2D_BWD),
                   ("dilated_spatial_backward_kernel_im2col", CONV2D_BWD),
                   ("cuboid_im2col", CONV3D_FWD), ("cuboid_mec", CONV3D_FWD), ("cuboid_naive", CONV3D_FWD),
                   ("cuboid_backward_input_im2col", CONV3D_BWD), ("cuboid_backward_kernel_im2col", CONV3D_BWD),
                   ("cuboid_backward_input_mec", CONV3D_BWD), ("cuboid_backward_kernel_mec", CONV3D_BWD),
                   ("cuboid_backward_input_naive", CONV3D_BWD), ("cuboid_backward_kernel_naive", CONV3D_BWD),
                   ("dilated_cuboid_im2col", DCONV3D_FWD),
                   ("dilated_cuboid_backward_input_im2col", DCONV3D_BWD),
                   ("dilated_cuboid_backward_kernel_im2col", DCONV3D_BWD)):
        DROPINS[f"owl_amd_{_n}_ndarray_conv_{_v}"] = _k
    DROPINS[f"owl_amd_{_n}_ndarray_upsampling_spatial_backward"] = UPSAMPLE
    DROPINS[f"owl_amd_{_n}_ndarray_upsampling_spatial"] = UPSAMPLE
    for _v, _k in (("maxpool_spatial", POOL2D_FWD), ("avgpool_spatial", POOL2D_FWD),
                   ("maxpool_spatial_backward", POOL2D_BWD), ("avgpool_spatial_backward", POOL2D_BWD),
                   ("maxpool_spatial_arg", POOL2D_ARG), ("maxpool_cuboid", POOL3D_FWD),
                   ("avgpool_cuboid", POOL3D_FWD), ("maxpool_cuboid_backward", POOL3D_BWD),
                   ("avgpool_cuboid_backward", POOL3D_BWD)):
        DROPINS[f"owl_amd_{_n}_ndarray_{_v}"] = _k


def lib():
    """Load the library (raises OwlAmdError if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OwlAmdError(ERR_UNSUPPORTED,
                              f"{LIB_PATH} is not built (run `python -m owl_amd.build`)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        cd, pd = ctypes.POINTER(ConvDesc), ctypes.POINTER(PoolDesc)
        sig = {
            "owl_amd_init": ([i32], i32),
            "owl_amd_last_error": ([], ctypes.c_char_p),
            "owl_amd_cache_stats": ([ctypes.POINTER(i64), ctypes.POINTER(i64)], i32),
            "owl_amd_dropin_graphs": ([i32], i32),
            "owl_amd_dropin_graph_stats": ([ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64)], i32),
            "owl_amd_cache_release": ([], None),
            "owl_amd_complex_supported": ([], i32),
            "owl_amd_kernel_timing": ([i32], i32),
            "owl_amd_kernel_time": ([i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64),
                                     ctypes.POINTER(ctypes.c_char_p)], i32),
            "owl_amd_conv_fwd": ([cd, i32, i32, vp, vp, vp, vp], i32),
            "owl_amd_conv_fwd_bias_act": ([cd, i32, i32, vp, vp, vp, i32, vp, vp, vp], i32),
            "owl_amd_conv_fwd_bias_act_pool": ([cd, i32, i32, vp, vp, vp, i32, vp, vp, vp, vp], i32),
            "owl_amd_bias_act_bwd": ([i32, i32, i64, i64, vp, vp, vp, vp, vp], i32),
            "owl_amd_maxpool_bias_act_bwd": ([i32, i32, i64, i64, i64, i64, vp, vp, vp, vp, vp, vp], i32),
            "owl_amd_upsampling2d_fwd": ([i32] + [i64] * 8 + [vp, vp, vp], i32),
            "owl_amd_upsampling2d_bwd": ([i32] + [i64] * 8 + [vp, vp, vp], i32),
            "owl_amd_conv_bwd_input": ([cd, i32, i32, vp, vp, vp, vp], i32),
            "owl_amd_conv_bwd_kernel_workspace": ([cd, i32], ctypes.c_size_t),
            "owl_amd_conv_bwd_kernel": ([cd, i32, i32, vp, vp, vp, vp, ctypes.c_size_t, vp], i32),
            "owl_amd_conv_bwd_kernel_noalloc": ([cd, i32, i32, vp, vp, vp, vp], i32),
            "owl_amd_conv_bwd": ([cd, i32, i32, i32, vp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp], i32),
            "owl_amd_pool_fwd": ([pd, i32, i32, vp, vp, vp], i32),
            "owl_amd_pool_argmax": ([pd, i32, vp, vp, vp, vp], i32),
            "owl_amd_pool_bwd": ([pd, i32, i32, vp, vp, vp, vp], i32),
            "owl_amd_conv_fwd_fp8": ([cd, i32, vp, vp, vp, vp, vp, vp], i32),
            "owl_amd_fp8_quantize": ([i32, vp, i64, vp, vp, vp], i32),
            "owl_amd_fp8_encode": ([vp, i64, vp, vp], i32),
            "owl_amd_conv_group_create": ([ctypes.POINTER(vp), i32, cd, i32, i32], i32),
            "owl_amd_conv_group_autotune": ([vp, vp, vp, vp, vp], i32),
            "owl_amd_conv_group_run": ([vp, vp, vp, vp, vp], i32),
            "owl_amd_conv_group_capture": ([vp, vp, vp, vp, vp], i32),
            "owl_amd_conv_group_replay": ([vp, vp], i32),
            "owl_amd_conv_group_info": ([vp, ctypes.POINTER(i32), ctypes.POINTER(i64)], i32),
            "owl_amd_conv_group_plan": ([vp, ctypes.POINTER(i32), ctypes.POINTER(i32)], i32),
            "owl_amd_conv_group_set_plan": ([vp, ctypes.POINTER(i32), ctypes.POINTER(i32)], i32),
            "owl_amd_conv_group_destroy": ([vp], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = res
        for name, kinds in DROPINS.items():
            f = getattr(L, name)
            f.argtypes = [vp if k == "p" else i64 for k in kinds]
            f.restype = i32
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise OwlAmdError(rc, lib().owl_amd_last_error().decode(errors="replace"))
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def kernel_timing(enable):
    """Start (clearing the totals) or stop recording kernel times."""
    check(lib().owl_amd_kernel_timing(1 if enable else 0))


def kernel_time(kclass):
    """(total ms, launches, kernel name) recorded for one kernel class
    (KT_FWD, KT_BWD_INPUT, KT_BWD_KERNEL, KT_AUX); waits for the events."""
    ms, n, name = ctypes.c_double(), ctypes.c_int64(), ctypes.c_char_p()
    check(lib().owl_amd_kernel_time(kclass, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(name)))
    return ms.value, n.value, (name.value or b"").decode()
