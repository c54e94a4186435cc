"""Pins for the oracle beyond Owl's golden vectors: an independent brute-force
convolution, the adjoint identities that tie the three passes together, the
1x1 == GEMM special case, Owl's FP32 CPU algorithm (im2col + OpenBLAS SGEMM)
and the pooling semantics (first strict max, NaN never wins, flat int64
argmax, gradient of an all-NaN window to flat index 0)."""
import itertools

import numpy as np
import pytest

import oracle
import oracle_ndarray as on
from owl_geometry import SAME, VALID, dilated_extent as k_eff, lead_pad as c_pad, out_dim as conv_out_dim
from owl_geometry import transpose_dim, wrapper_pads_2d


def brute_conv(x, w, stride, rate, pads, out_sp):
    """Direct definition with explicit zero padding, written independently of
    the oracle: y[b,o,co] = sum_{t,ci} xpad[b, o*s + t*r, ci] w[t, ci, co]."""
    nd = x.ndim - 2
    B, Ci, Co = x.shape[0], x.shape[-1], w.shape[-1]
    ksp = w.shape[:nd]
    big = [max(n + p, o * s + (k - 1) * r + 1) for n, p, o, s, k, r in
           zip(x.shape[1:-1], pads, out_sp, stride, ksp, rate)]
    xp = np.zeros([B, *big, Ci])
    sl = tuple(slice(p, p + n) for p, n in zip(pads, x.shape[1:-1]))
    xp[(slice(None),) + sl] = x
    y = np.zeros([B, *out_sp, Co])
    for o in itertools.product(*[range(n) for n in out_sp]):
        for t in itertools.product(*[range(k) for k in ksp]):
            idx = tuple(oo * s + tt * r for oo, s, tt, r in zip(o, stride, t, rate))
            y[(slice(None),) + o] += xp[(slice(None),) + idx] @ w[t]
    return y


def random_problem(rng, nd):
    sp = [int(rng.integers(3, 8)) for _ in range(nd)]
    ksp = [int(rng.integers(1, 4)) for _ in range(nd)]
    stride = [int(rng.integers(1, 3)) for _ in range(nd)]
    rate = [int(rng.integers(1, 3)) for _ in range(nd)]
    pad = [SAME, VALID][int(rng.integers(0, 2))]
    B, Ci, Co = int(rng.integers(1, 3)), int(rng.integers(1, 5)), int(rng.integers(1, 5))
    for i in range(nd):
        while conv_out_dim(pad, sp[i], k_eff(ksp[i], rate[i]), stride[i]) <= 0:
            sp[i] += 2
    x = rng.standard_normal([B, *sp, Ci])
    w = rng.standard_normal([*ksp, Ci, Co])
    return x, w, stride, rate, pad


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("nd", [2, 3])
def test_oracle_forward_matches_bruteforce(seed, nd):
    rng = np.random.default_rng(seed * 7 + nd)
    x, w, stride, rate, pad = random_problem(rng, nd)
    y = on.dilated_conv2d(x, w, stride, rate, pad)
    sp = x.shape[1:-1]
    out = y.shape[1:-1]
    pads = [c_pad(s, o, k_eff(k, r), n) for n, k, o, s, r in zip(sp, w.shape[:nd], out, stride, rate)] \
        if pad == SAME else [0] * nd
    ref = brute_conv(x, w, stride, rate, pads, out)
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("nd", [2, 3])
def test_oracle_adjoint_identities(seed, nd):
    """<conv(x,w), g> = <x, dgrad(w,g)> = <w, wgrad(x,g)>: a dropped tap, a
    wrong sign of the pad or a transposed kernel index breaks one of them."""
    rng = np.random.default_rng(100 + seed * 3 + nd)
    x, w, stride, rate, pad = random_problem(rng, nd)
    y = on.dilated_conv2d(x, w, stride, rate, pad)
    g = rng.standard_normal(y.shape)
    dx = on.dilated_conv2d_backward_input(x, w, stride, rate, g)
    dw = on.dilated_conv2d_backward_kernel(x, w, stride, rate, g)
    lhs = float((y * g).sum())
    assert abs(lhs - float((x * dx).sum())) <= 1e-10 * (1 + abs(lhs))
    assert abs(lhs - float((w * dw).sum())) <= 1e-10 * (1 + abs(lhs))


def test_oracle_1x1_is_gemm():
    rng = np.random.default_rng(5)
    x = rng.standard_normal([3, 5, 4, 7])
    w = rng.standard_normal([1, 1, 7, 6])
    y = on.conv2d(x, w, [1, 1], VALID)
    np.testing.assert_allclose(y, (x.reshape(-1, 7) @ w[0, 0]).reshape(3, 5, 4, 6), rtol=1e-13, atol=1e-13)
    g = rng.standard_normal(y.shape)
    np.testing.assert_allclose(on.conv2d_backward_input(x, w, [1, 1], g),
                               (g.reshape(-1, 6) @ w[0, 0].T).reshape(x.shape), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(on.conv2d_backward_kernel(x, w, [1, 1], g)[0, 0],
                               x.reshape(-1, 7).T @ g.reshape(-1, 6), rtol=1e-12, atol=1e-12)


def test_oracle_wgrad_matches_finite_differences():
    """The kernel gradient of L = sum(conv(x, w) * g) by central differences."""
    rng = np.random.default_rng(9)
    x = rng.standard_normal([2, 6, 5, 3])
    w = rng.standard_normal([3, 2, 3, 4])
    y = on.conv2d(x, w, [2, 1], SAME)
    g = rng.standard_normal(y.shape)
    dw = on.conv2d_backward_kernel(x, w, [2, 1], g)
    eps = 1e-4
    for idx in [(0, 0, 0, 0), (2, 1, 2, 3), (1, 0, 1, 2)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps
        wm[idx] -= eps
        fd = ((on.conv2d(x, wp, [2, 1], SAME) * g).sum() - (on.conv2d(x, wm, [2, 1], SAME) * g).sum()) / (2 * eps)
        assert abs(fd - dw[idx]) < 1e-8 * (1 + abs(fd))


@pytest.mark.parametrize("shape", [([4, 9, 9, 16], [3, 3, 16, 24], [1, 1], SAME),
                                   ([2, 11, 7, 8], [3, 1, 8, 12], [2, 2], VALID),
                                   ([2, 8, 8, 256], [1, 1, 256, 64], [2, 2], SAME)])
def test_owl_fp32_algorithm_matches_fp64(shape):
    """Owl's CPU algorithm (im2col + one SGEMM, owl_ndarray_conv_impl.h:297-338)
    restated over the host OpenBLAS, against the FP64 oracle."""
    xs, ks, stride, pad = shape
    rng = np.random.default_rng(1)
    x = rng.standard_normal(xs).astype(np.float32)
    w = rng.standard_normal(ks).astype(np.float32)
    y64 = on.conv2d(x.astype(np.float64), w.astype(np.float64), stride, pad)
    sp = x.shape[1:3]
    out = y64.shape[1:3]
    pads = [c_pad(s, o, k, n) for n, k, o, s in zip(sp, ks[:2], out, stride)] if pad == SAME else [0, 0]
    g = oracle.conv_geom(xs, ks, out, stride, [1, 1], pads)
    y32 = oracle.conv_fwd_owl_f32(g, x, w).reshape(y64.shape)
    assert np.linalg.norm(y32 - y64) / np.linalg.norm(y64) < 1e-5
    dy = rng.standard_normal(y64.shape).astype(np.float32)
    dx64 = on.conv2d_backward_input(x.astype(np.float64), w.astype(np.float64), stride, dy.astype(np.float64))
    dx32 = oracle.conv_bwd_input_owl_f32(g, w, dy).reshape(x.shape)
    assert np.linalg.norm(dx32 - dx64) / np.linalg.norm(dx64) < 1e-5
    dw64 = on.conv2d_backward_kernel(x.astype(np.float64), w.astype(np.float64), stride, dy.astype(np.float64))
    dw32 = oracle.conv_bwd_kernel_owl_f32(g, x, dy).reshape(w.shape)
    assert np.linalg.norm(dw32 - dw64) / np.linalg.norm(dw64) < 1e-5


# ------------------------------------------------------------------ pooling
def brute_pool(x, k, s, pads, out, is_max):
    y = np.zeros([x.shape[0], *out, x.shape[-1]], x.dtype)
    for o in itertools.product(*[range(n) for n in out]):
        sl = []
        for oo, ss, p, kk, n in zip(o, s, pads, k, x.shape[1:-1]):
            lo, hi = max(oo * ss - p, 0), min(oo * ss - p + kk, n)
            sl.append(slice(lo, hi))
        win = x[(slice(None),) + tuple(sl)].reshape(x.shape[0], -1, x.shape[-1])
        y[(slice(None),) + o] = win.max(axis=1) if is_max else win.sum(axis=1, dtype=np.float64) / win.shape[1]
    return y


@pytest.mark.parametrize("seed", range(6))
def test_oracle_pool_matches_bruteforce(seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal([2, 7, 6, 3])
    k, s = [int(rng.integers(1, 4)), int(rng.integers(1, 4))], [int(rng.integers(1, 3)), int(rng.integers(1, 3))]
    for pad in (SAME, VALID):
        ym = on.max_pool2d(x, k, s, pad)
        ya = on.avg_pool2d(x, k, s, pad)
        out = ym.shape[1:3]
        pads = [c_pad(ss, o, kk, n) for n, kk, o, ss in zip(x.shape[1:3], k, out, s)] if pad == SAME else [0, 0]
        np.testing.assert_array_equal(ym, brute_pool(x, k, s, pads, out, True))
        np.testing.assert_allclose(ya, brute_pool(x, k, s, pads, out, False), rtol=1e-14)
        y2, arg = on.max_pool2d_argmax(x, k, s, pad)
        np.testing.assert_array_equal(y2, ym)
        np.testing.assert_array_equal(x.ravel()[arg.ravel()], ym.ravel())  # argmax is a flat index
        dy = rng.standard_normal(ym.shape)
        dx = on.max_pool2d_backward(pad, x, k, s, dy)
        assert abs(dx.sum() - dy.sum()) < 1e-12 * (1 + abs(dy).sum())  # mass conservation
        dxa = on.avg_pool2d_backward(pad, x, k, s, dy)
        assert abs(dxa.sum() - dy.sum()) < 1e-12 * (1 + abs(dy).sum())


def test_oracle_pool_ties_nan_and_orphans():
    """First strict max wins, NaN never wins, -1 argmax for an all-NaN window,
    and max backward sends an all-NaN window's gradient to flat index 0
    (owl_ndarray_pool_impl.h:165-195, :516-536)."""
    x = np.array([1, 3, 3, 2,
                  np.nan, np.nan, 5, 0,
                  np.nan, np.nan, 0, 0,
                  np.nan, np.nan, 0, 7], np.float32).reshape(1, 4, 4, 1)
    y, arg = on.max_pool2d_argmax(x, [2, 2], [2, 2], VALID)
    assert y.ravel().tolist()[1:] == [5, -np.inf, 7] and y.ravel()[0] == 3
    assert arg.ravel().tolist() == [1, 6, -1, 15]
    dy = np.array([1, 2, 4, 8], np.float32).reshape(1, 2, 2, 1)
    dx = on.max_pool2d_backward(VALID, x, [2, 2], [2, 2], dy).ravel()
    assert dx[1] == 1 and dx[6] == 2 and dx[15] == 8 and dx[0] == 4
    assert dx.sum() == 15


def test_geometry_matches_product():
    """The product's shape / pad formulas (owl_amd.shapes, Owl's float form)
    against the test side's integer restatement, over every small size,
    kernel, stride and rate; plus the textbook special cases: SAME at stride 1
    keeps the size and pads (k_eff - 1) / 2 in front for odd extents; VALID at
    stride 1 gives n - k_eff + 1 and no pad."""
    from owl_amd import shapes as P
    for n in range(1, 40):
        for k in range(1, 8):
            for s in range(1, 6):
                for r in range(1, 4):
                    ke = k_eff(k, r)
                    assert P.k_eff(k, r) == ke
                    for pad in (SAME, VALID):
                        o = conv_out_dim(pad, n, ke, s)
                        assert P.conv_out_dim(pad, n, ke, s) == o, (pad, n, ke, s)
                        assert P.transpose_out_dim(pad, n, k, s) == transpose_dim(pad, n, k, s)
                        if o > 0:
                            assert P.c_pad(s, o, ke, n) == c_pad(s, o, ke, n), (n, ke, s, o)
                    assert P.calc_conv2d_padding(n, n + 1, k, ke, conv_out_dim(SAME, n, k, s),
                                                 conv_out_dim(SAME, n + 1, ke, s + 1), s + 1, s) == \
                        wrapper_pads_2d(n, n + 1, k, ke, conv_out_dim(SAME, n, k, s),
                                        conv_out_dim(SAME, n + 1, ke, s + 1), s + 1, s)
                    if s == 1:
                        assert conv_out_dim(SAME, n, ke, 1) == n
                        if ke % 2 == 1:
                            assert c_pad(1, n, ke, n) == (ke - 1) // 2
                        if n >= ke:
                            assert conv_out_dim(VALID, n, ke, 1) == n - ke + 1
                            assert c_pad(1, n - ke + 1, ke, n) == 0
