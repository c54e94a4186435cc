"""Replays Owl's conv / pool golden tests against a backend.

``tests/golden/owl_vectors.json`` holds each test's call and expected values
(extracted from the reference's test/*.ml by extract_owl_vectors.py).  This
module restates how each OCaml test helper builds its inputs and which
Dense.Ndarray function it calls, then evaluates it with a backend that has
``owl_amd.ndarray``'s signatures (the oracle mirror or the GPU library).
"""
import json
import os

import numpy as np

from owl_amd.shapes import SAME, VALID

HERE = os.path.dirname(os.path.abspath(__file__))


def load_cases():
    with open(os.path.join(HERE, "golden", "owl_vectors.json")) as f:
        return json.load(f)["cases"]


def case_id(c):
    return f"{os.path.basename(c['file'])[:-3]}:{c['module']}.{c['fun']}"


def seq(shape, a, dtype):
    """N.sequential ~a (owl_dense_ndarray_generic.ml:1471-1484): a, a+1, ... row-major."""
    n = int(np.prod(shape))
    return (np.arange(n, dtype=np.float64) + a).astype(dtype).reshape(shape)


def ones(shape, dtype):
    return np.ones(shape, dtype=dtype)


def seq_div(shape, dtype):
    """N.sequential ~a:1. then div_scalar by the element count, in the array's dtype
    (test/unit_conv3d_generic.ml:16-23)."""
    s = seq(shape, 1.0, dtype)
    return (s / np.asarray(s.size, dtype=dtype)).astype(dtype)


def run_case(c, nd, dtype=np.float32):
    """Evaluate golden case ``c`` with backend module ``nd``; returns the result array."""
    h = c["helper"]
    ins, ks, third, pad = c["input_shape"], c["kernel_shape"], c["third"], c["padding"]
    S = lambda shp, a=1.0: seq(shp, a, dtype)
    # conv2d / conv3d (test/unit_conv2d_generic.ml:16-45, unit_conv3d_generic.ml:16-40)
    if h in ("test_conv2d", "test_conv3d"):
        x, w = (S(ins), S(ks)) if h == "test_conv2d" else (seq_div(ins, dtype), seq_div(ks, dtype))
        f = nd.conv2d if h == "test_conv2d" else nd.conv3d
        return f(x, w, third, pad)
    if h in ("test_conv2d_back_input", "test_conv2d_back_kernel", "test_conv2d_back_kernel_02",
             "test_conv3d_back_input", "test_conv3d_back_kernel"):
        is3 = "3d" in h
        if h == "test_conv2d_back_kernel_02":
            x, w = ones(ins, dtype), ones(ks, dtype)
        else:
            x, w = S(ins), S(ks)
        y = (nd.conv3d if is3 else nd.conv2d)(x, w, third, pad)
        if h == "test_conv2d_back_kernel_02":
            dy = ones(y.shape, dtype)
        else:
            dy = S(y.shape, 0.0 if is3 else 1.0)  # conv3d tests use ~a:0.
        if "input" in h:
            return (nd.conv3d_backward_input if is3 else nd.conv2d_backward_input)(x, w, third, dy)
        return (nd.conv3d_backward_kernel if is3 else nd.conv2d_backward_kernel)(x, w, third, dy)
    # pooling (test/unit_pool2d_generic.ml:16-41, unit_pool3d_generic.ml:16-41)
    if h in ("test_maxpool2d", "test_avgpool2d", "test_maxpool3d", "test_avgpool3d"):
        f = getattr(nd, {"test_maxpool2d": "max_pool2d", "test_avgpool2d": "avg_pool2d",
                         "test_maxpool3d": "max_pool3d", "test_avgpool3d": "avg_pool3d"}[h])
        return f(S(ins), ks, third, pad)
    if h in ("test_maxpool2d_back", "test_avgpool2d_back", "test_maxpool3d_back", "test_avgpool3d_back"):
        base = h[len("test_"):-len("_back")]
        name = {"maxpool2d": "max_pool2d", "avgpool2d": "avg_pool2d",
                "maxpool3d": "max_pool3d", "avgpool3d": "avg_pool3d"}[base]
        x = S(ins)
        y = getattr(nd, name)(x, ks, third, pad)
        dy = S(y.shape)
        return getattr(nd, name + "_backward")(pad, x, ks, third, dy)
    # dilated (test/unit_dilated_conv2d.ml:14-35, unit_dilated_conv3d.ml:14-35)
    if h.startswith("test_dilated_conv"):
        is3 = "conv3d" in h
        stride = c.get("stride", [1, 1, 1] if is3 else [1, 1])
        rate = third
        x, w = S(ins), S(ks)
        f = nd.dilated_conv3d if is3 else nd.dilated_conv2d
        if h.endswith("_bi") or h.endswith("_bk"):
            y = f(x, w, stride, rate, pad)
            a0 = 0.0 if (is3 and h.endswith("_bi")) else 1.0
            dy = S(y.shape, a0)
            g = getattr(nd, ("dilated_conv3d" if is3 else "dilated_conv2d") +
                        ("_backward_input" if h.endswith("_bi") else "_backward_kernel"))
            return g(x, w, stride, rate, dy)
        return f(x, w, stride, rate, pad)
    # transpose (test/unit_transpose_conv2d.ml:14-46, unit_transpose_conv3d.ml:14-40)
    if h.startswith("compute_trans_conv") or h == "test_forward":
        is3 = "conv3d" in h or c["file"].endswith("conv3d.ml")
        tf = nd.transpose_conv3d if is3 else nd.transpose_conv2d
        seqf = c.get("seq", False)
        x = ones(ins, dtype)
        if h == "test_forward" or h.endswith("conv2d") or h.endswith("conv3d"):
            w = S(ks) if seqf else ones(ks, dtype)
            return tf(x, w, third, pad)
        if h.endswith("_bi"):
            w = S(ks) if seqf else ones(ks, dtype)
            y = tf(x, w, third, pad)
            dy = ones(y.shape, dtype)
            return (nd.transpose_conv3d_backward_input if is3 else nd.transpose_conv2d_backward_input)(
                x, w, third, dy)
        if h.endswith("_bk"):
            w = ones(ks, dtype)
            y = tf(x, w, third, pad)
            dy = S(y.shape) if seqf else ones(y.shape, dtype)
            return (nd.transpose_conv3d_backward_kernel if is3 else nd.transpose_conv2d_backward_kernel)(
                x, w, third, dy)
    raise KeyError(f"unknown helper {h}")


def transpose_magic_expected(c, shape):
    """Expected output of test/unit_transpose_conv2d.ml:49-85 (test_forward):
    a closed form in three "magic numbers" for all-ones input and kernel."""
    stride, pad, magic = c["third"], c["padding"], c["magic"]
    ph = 0 if pad == SAME else stride[0] - 1
    pw = 0 if pad == SAME else stride[1] - 1
    e = np.zeros(shape, dtype=np.float64)
    for n in range(shape[0]):
        for k in range(shape[3]):
            for w in range(pw, shape[2] - pw):
                for h in range(ph, shape[1] - ph):
                    t = magic[0]
                    h_in = h % stride[0] == 0 and ph < h < shape[1] - 1 - ph
                    w_in = w % stride[1] == 0 and pw < w < shape[2] - 1 - pw
                    addon = magic[1] if (h_in and w_in) else (magic[2] if (h_in or w_in) else 0.0)
                    e[n, h, w, k] = t + addon
            if pad == VALID:
                e[n, :, 0, k] = e[n, :, 1, k]
                e[n, :, -1, k] = e[n, :, -2, k]
                e[n, 0, :, k] = e[n, 1, :, k]
                e[n, -1, :, k] = e[n, -2, :, k]
    return e


def expected_of(c, result_shape):
    if c["helper"] == "test_forward":
        return transpose_magic_expected(c, result_shape)
    return np.asarray(c["expected"], dtype=np.float64).reshape(result_shape)
