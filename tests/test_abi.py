"""The C-ABI library loads and exports every entry point include/owl_amd.h
declares; argument errors come back as status codes (no exit, no GPU call)."""
import ctypes
import os
import re
import subprocess

import pytest

from owl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "owl_amd.h")


def declared_functions():
    out = subprocess.run(["gcc", "-E", "-P", "-x", "c", HEADER], capture_output=True, text=True, check=True).stdout
    return sorted(set(re.findall(r"\b(owl_amd_\w+)\s*\(", out)))


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_library_exports_every_declared_symbol():
    decl = declared_functions()
    assert len(decl) >= 80
    missing = [d for d in decl if d not in exported()]
    assert not missing, missing


def test_dropin_table_matches_header():
    decl = set(declared_functions())
    assert set(_lib.DROPINS) <= decl
    assert {d for d in decl if "_ndarray_" in d} == set(_lib.DROPINS)
    for name in ("owl_amd_conv_group_create", "owl_amd_conv_group_run", "owl_amd_conv_group_replay"):
        assert name in decl


def test_group_plan_calls_check_arguments():
    L = _lib.lib()
    assert L.owl_amd_conv_group_set_plan(None, None, None) != 0
    assert L.owl_amd_conv_group_plan(None, None, None) != 0


def test_library_loads_and_binds():
    L = _lib.lib()
    for name in _lib.DROPINS:
        assert getattr(L, name).restype == ctypes.c_int
    assert L.owl_amd_complex_supported() == 1


def _desc(**kw):
    d = dict(B=1, H=4, W=4, D=1, Ci=2, kh=3, kw=3, kd=1, Co=2, Ho=4, Wo=4, Do=1,
             sh=1, sw=1, sd=1, rh=1, rw=1, rd=1, ph=1, pw=1, pd=0)
    d.update(kw)
    return _lib.ConvDesc(**d)


@pytest.mark.parametrize("bad", [dict(B=-1), dict(sh=0), dict(rw=0), dict(kh=0), dict(Ci=1 << 33)])
def test_bad_conv_arguments_return_status(bad):
    L = _lib.lib()
    d = _desc(**bad)
    buf = ctypes.c_void_p(16)  # never dereferenced: the check fails first
    assert L.owl_amd_conv_fwd(ctypes.byref(d), _lib.F32, _lib.F32, buf, buf, buf, None) == _lib.ERR_ARG
    assert L.owl_amd_conv_bwd_input(ctypes.byref(d), _lib.F32, _lib.F32, buf, buf, buf, None) == _lib.ERR_ARG
    assert L.owl_amd_conv_bwd_kernel(ctypes.byref(d), _lib.F32, _lib.F32, buf, buf, buf, None, 0, None) \
        == _lib.ERR_ARG
    assert L.owl_amd_last_error()


def test_null_descriptor_and_pointers():
    L = _lib.lib()
    assert L.owl_amd_conv_fwd(None, 0, 0, None, None, None, None) == _lib.ERR_ARG
    d = _desc()
    assert L.owl_amd_conv_fwd(ctypes.byref(d), _lib.F32, _lib.F32, None, None, None, None) == _lib.ERR_ARG


def test_dropin_rejects_zero_stride_without_touching_the_gpu():
    L = _lib.lib()
    p = ctypes.c_void_p(16)
    rc = L.owl_amd_float32_ndarray_conv_spatial(p, p, p, 1, 4, 4, 1, 3, 3, 4, 4, 1, 0, 1, 0, 1, 1)
    assert rc == _lib.ERR_ARG
    rc = L.owl_amd_float64_ndarray_maxpool_spatial(p, p, 1, 4, 4, 1, 2, 2, 2, 2, 2, 0, 1, 1, 1)
    assert rc == _lib.ERR_ARG


def test_wrapper_shape_errors_mirror_owl():
    import numpy as np
    from owl_amd import ndarray as nd
    with pytest.raises(nd.InvalidArgument):
        nd.conv2d(np.zeros([1, 4, 4, 3], np.float32), np.zeros([3, 3, 2, 4], np.float32), [1, 1])
    with pytest.raises(nd.InvalidArgument):
        nd.conv2d(np.zeros([1, 4, 4, 3], np.float32), np.zeros([3, 3, 3, 4], np.float32), [1, 1, 1])
    with pytest.raises(nd.InvalidArgument):
        nd.max_pool2d(np.zeros([1, 4, 4], np.float32), [2, 2], [2, 2])
