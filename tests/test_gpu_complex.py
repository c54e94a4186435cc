"""Owl's complex kinds (complex32 = pairs of float32, complex64 = pairs of
float64) on the GPU: the drop-ins (numpy complex arrays) and the device entry
points (torch complex tensors).

Convolutions: Owl's complex stubs run the real algorithm with complex GEMMs
that transpose without conjugating (owl_ndarray_conv_stub.c:91-122), so with
x = xr + i xi and w = wr + i wi each pass is fixed by four real ones:
  conv(x, w)          = conv(xr, wr) - conv(xi, wi) + i (conv(xr, wi) + conv(xi, wr))
  bwd_input(w, dy)    = bi(wr, dyr) - bi(wi, dyi) + i (bi(wi, dyr) + bi(wr, dyi))
  bwd_kernel(x, dy)   = bk(xr, dyr) - bk(xi, dyi) + i (bk(xr, dyi) + bk(xi, dyr))
which are evaluated with the FP64 oracle; tolerance rel-L2 1e-5 (complex32)
and 1e-12 (complex64).  The FP64 cases also check the decomposition itself
against a direct complex loop on a tiny problem.

Average pooling and upsampling are componentwise (Owl's complex sum and its
division by the integer count act on each part), so they equal the real
results on the interleaved parts bit for bit.  Complex max pooling is not
provided: Owl's starts from -inf + 0i, whose modulus is +inf, so its
modulus-then-argument comparison never replaces it; the call reports
unsupported."""
import numpy as np
import pytest

import oracle
import oracle_ndarray as on
from owl_amd import _lib
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")

DEV = "cuda:0"
TOL = {np.complex64: 1e-5, np.complex128: 1e-12}


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def crand(rng, shape, dt, scale=1.0):
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * scale).astype(dt)


def cfwd(f, x, w):
    xr, xi, wr, wi = x.real, x.imag, w.real, w.imag
    return (f(xr, wr) - f(xi, wi)) + 1j * (f(xr, wi) + f(xi, wr))


def both(fn, *args):
    """The call on numpy arrays (drop-ins) and on device tensors."""
    h = fn(*args)
    d = fn(*[torch.from_numpy(a).to(DEV) if isinstance(a, np.ndarray) else a for a in args]).cpu().numpy()
    return h, d


CASES = [  # (x shape, kernel shape, stride, padding, rate)
    ([2, 9, 8, 3], [3, 3, 3, 5], [1, 1], SAME, None),
    ([2, 11, 10, 4], [3, 2, 4, 6], [2, 1], VALID, None),
    ([1, 12, 12, 8], [3, 3, 8, 8], [1, 1], SAME, [2, 2]),
    ([2, 5, 6, 4, 3], [2, 3, 2, 3, 4], [1, 2, 1], SAME, None),
    ([2, 9, 10, 6], [1, 1, 6, 4], [2, 3], SAME, None),   # 1x1 with strides: subsampled grid + scatter
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_complex_conv_passes(case, dt):
    xs, ks, stride, pad, rate = CASES[case]
    rng = np.random.default_rng(60 + case)
    x = crand(rng, xs, dt)
    w = crand(rng, ks, dt, 1 / np.sqrt(np.prod(ks[:-1])))
    three = len(xs) == 5
    if rate is not None:
        fwd = lambda a, b: nd.dilated_conv2d(a, b, stride, rate, pad)
        ofwd = lambda a, b: on.dilated_conv2d(a, b, stride, rate, pad)
        bi = lambda a, b, g: nd.dilated_conv2d_backward_input(a, b, stride, rate, g)
        obi = lambda a, b, g: on.dilated_conv2d_backward_input(a, b, stride, rate, g)
        bk = lambda a, b, g: nd.dilated_conv2d_backward_kernel(a, b, stride, rate, g)
        obk = lambda a, b, g: on.dilated_conv2d_backward_kernel(a, b, stride, rate, g)
    else:
        c = nd.conv3d if three else nd.conv2d
        fwd = lambda a, b: c(a, b, stride, pad)
        ofwd = lambda a, b: on.conv2d(a, b, stride, pad)
        f_bi = nd.conv3d_backward_input if three else nd.conv2d_backward_input
        f_bk = nd.conv3d_backward_kernel if three else nd.conv2d_backward_kernel
        bi = lambda a, b, g: f_bi(a, b, stride, g)
        bk = lambda a, b, g: f_bk(a, b, stride, g)
        obi = lambda a, b, g: on.conv2d_backward_input(a, b, stride, g)
        obk = lambda a, b, g: on.conv2d_backward_kernel(a, b, stride, g)
    x64, w64 = x.astype(np.complex128), w.astype(np.complex128)
    ref = cfwd(ofwd, x64, w64)
    for got in both(fwd, x, w):
        assert got.dtype == dt and got.shape == ref.shape
        assert rel(got, ref) < TOL[dt], rel(got, ref)
    dy = crand(rng, ref.shape, dt)
    dy64 = dy.astype(np.complex128)
    xr, xi, wr, wi, gr, gi = x64.real, x64.imag, w64.real, w64.imag, dy64.real, dy64.imag
    ref_dx = (obi(xr, wr, gr) - obi(xr, wi, gi)) + 1j * (obi(xr, wi, gr) + obi(xr, wr, gi))
    ref_dw = (obk(xr, wr, gr) - obk(xi, wr, gi)) + 1j * (obk(xr, wr, gi) + obk(xi, wr, gr))
    for got in both(bi, x, w, dy):
        assert rel(got, ref_dx) < TOL[dt], rel(got, ref_dx)
    for got in both(bk, x, w, dy):
        assert rel(got, ref_dw) < TOL[dt], rel(got, ref_dw)


@pytest.mark.gpu
def test_complex_decomposition_against_direct_loops():
    # the identities above, against plain complex loops on a tiny problem
    rng = np.random.default_rng(5)
    x = crand(rng, [1, 4, 5, 2], np.complex128)
    w = crand(rng, [2, 3, 2, 3], np.complex128)
    y = nd.conv2d(x, w, [1, 1], VALID)
    ref = np.zeros(y.shape, np.complex128)
    for i in range(y.shape[1]):
        for j in range(y.shape[2]):
            ref[0, i, j] = np.einsum("abq,abqo->o", x[0, i:i + 2, j:j + 3], w)
    assert rel(y, ref) < 1e-13
    dy = crand(rng, y.shape, np.complex128)
    dx = nd.conv2d_backward_input(x, w, [1, 1], dy)
    dw = nd.conv2d_backward_kernel(x, w, [1, 1], dy)
    rdx = np.zeros(x.shape, np.complex128)
    rdw = np.zeros(w.shape, np.complex128)
    for i in range(y.shape[1]):
        for j in range(y.shape[2]):
            rdx[0, i:i + 2, j:j + 3] += np.einsum("o,abqo->abq", dy[0, i, j], w)  # no conjugate
            rdw += np.einsum("abq,o->abqo", x[0, i:i + 2, j:j + 3], dy[0, i, j])
    assert rel(dx, rdx) < 1e-13 and rel(dw, rdw) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_complex_transpose_conv(dt):
    rng = np.random.default_rng(11)
    x = crand(rng, [2, 5, 6, 4], dt)
    w = crand(rng, [3, 3, 4, 3], dt, 0.2)  # input channels in the 3rd dimension, as Owl checks
    x64, w64 = x.astype(np.complex128), w.astype(np.complex128)
    ref = cfwd(lambda a, b: on.transpose_conv2d(a, b, [2, 2], SAME), x64, w64)
    for got in both(lambda a, b: nd.transpose_conv2d(a, b, [2, 2], SAME), x, w):
        assert rel(got, ref) < TOL[dt]


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_complex_avg_pool_and_upsampling_bit_exact(dt):
    rng = np.random.default_rng(12)
    x = crand(rng, [2, 9, 7, 5], dt)
    rdt = np.float32 if dt == np.complex64 else np.float64
    view = lambda a: a.view(rdt)  # [.., C] complex = [.., 2C] real
    for kern, stride, pad in (([2, 2], [2, 2], VALID), ([3, 2], [2, 1], SAME)):
        ref = on.avg_pool2d(view(x), kern, stride, pad).view(dt)
        for got in both(lambda a: nd.avg_pool2d(a, kern, stride, pad), x):
            assert np.array_equal(got, ref)
        dy = crand(rng, ref.shape, dt)
        rdx = on.avg_pool2d_backward(pad, view(x), kern, stride, view(dy)).view(dt)
        for got in both(lambda a, g: nd.avg_pool2d_backward(pad, a, kern, stride, g), x, dy):
            assert np.array_equal(got, rdx)
    up = nd.upsampling2d(x, [2, 3])
    assert np.array_equal(up, x.repeat(2, axis=1).repeat(3, axis=2))
    g = crand(rng, up.shape, dt)
    gx = nd.upsampling2d_backward(x, [2, 3], g)
    assert np.array_equal(gx, oracle.upsampling_bwd(view(g), list(view(x).shape), [2, 3]).view(dt))


@pytest.mark.gpu
def test_complex_max_pool_is_unsupported():
    x = np.ones([1, 4, 4, 2], np.complex64)
    with pytest.raises(_lib.OwlAmdError):
        nd.max_pool2d(x, [2, 2], [2, 2])
    with pytest.raises(_lib.OwlAmdError):
        nd.max_pool2d(torch.from_numpy(x).to(DEV), [2, 2], [2, 2])
    with pytest.raises(_lib.OwlAmdError):
        nd.max_pool2d_backward(VALID, x, [2, 2], [2, 2], np.ones([1, 2, 2, 2], np.complex64))
