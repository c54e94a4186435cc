"""The computation-graph engine on the GPU.

* ``eval_graph`` against the same operators called one by one through the
  ndarray wrappers: bit for bit in FP32 and FP64, both for the graph as built
  and after ``optimise`` (the fused conv epilogue adds the bias to the FP32
  accumulator and applies fmax, which is what the unfused kernels and
  torch's add / fmax do in FP32).
* ``capture_graph`` / ``replay_graph``: each replay, with new values assigned
  to the variables and eager work run between replays, equals an eager
  evaluation on the same values, bit for bit.
* BF16 / FP16: the graph as built equals the unfused operators run one by
  one, and the optimised graph the fused kernel followed by the same
  operators, bit for bit (the fusion itself rounds once instead of twice, so
  the two graphs differ from each other; the fused layer is within 1e-2 of
  its FP32 result)."""
import numpy as np
import pytest

from owl_amd import cgraph as G
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID

torch = pytest.importorskip("torch")

DEV = "cuda:0"


def build(dt, B=2):
    x = G.var_arr("x", [B, 16, 16, 32], dt)
    w1 = G.var_arr("w1", [3, 3, 32, 64], dt)
    b1 = G.var_arr("b1", [64], dt)
    w2 = G.var_arr("w2", [3, 3, 64, 64], dt)
    dy = G.var_arr("dy", [B, 8, 8, 64], dt)
    h = G.relu(G.add(G.conv2d(x, w1, [1, 1]), b1))
    p = G.max_pool2d(h, [2, 2], [2, 2])
    y = G.avg_pool2d(G.conv2d(p, w2, [1, 1]), [2, 2], [2, 2], VALID)
    dw = G.conv2d_backward_kernel(p, w2, [1, 1], dy)
    dp = G.conv2d_backward_input(p, w2, [1, 1], dy)
    dh = G.max_pool2d_backward(SAME, h, [2, 2], [2, 2], dp)
    up = G.upsampling2d(p, [2, 2])
    return [x, w1, b1, w2, dy], [y, dw, dh, up]


def values(dt, seed, B=2):
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    r = lambda *s, scale=1.0: (torch.randn(*s, device=DEV, generator=g) * scale).to(dt)
    return [r(B, 16, 16, 32), r(3, 3, 32, 64, scale=0.06), r(64, scale=0.3), r(3, 3, 64, 64, scale=0.04),
            r(B, 8, 8, 64)]


def eager(vals):
    x, w1, b1, w2, dy = vals
    h = torch.fmax(nd.conv2d(x, w1, [1, 1]) + b1, torch.zeros((), device=DEV, dtype=x.dtype))
    p = nd.max_pool2d(h, [2, 2], [2, 2])
    y = nd.avg_pool2d(nd.conv2d(p, w2, [1, 1]), [2, 2], [2, 2], VALID)
    dw = nd.conv2d_backward_kernel(p, w2, [1, 1], dy)
    dp = nd.conv2d_backward_input(p, w2, [1, 1], dy)
    dh = nd.max_pool2d_backward(SAME, h, [2, 2], [2, 2], dp)
    up = nd.upsampling2d(p, [2, 2])
    return [y, dw, dh, up]


def assign(ins, vals):
    for n, v in zip(ins, vals):
        G.assign_arr(n, v)


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
@pytest.mark.parametrize("opt", [False, True])
def test_eval_equals_eager(dt, opt):
    ins, outs = build(dt)
    g = G.make_graph(ins, outs, "cnn")
    if opt:
        G.optimise(g)
        assert "conv2d_bias_act" in [n.op for n in g.order]
    vals = values(dt, 1)
    assign(ins, vals)
    got = G.eval_graph(g)
    ref = eager(vals)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    # a second evaluation on new values reuses the same blocks
    ptrs = [t.data_ptr() for t in got]
    vals2 = values(dt, 2)
    assign(ins, vals2)
    got2 = G.eval_graph(g)
    assert [t.data_ptr() for t in got2] == ptrs
    for a, b in zip(got2, eager(vals2)):
        assert torch.equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("opt", [False, True])
def test_capture_replay(opt):
    dt = torch.float32
    ins, outs = build(dt, B=4)
    g = G.make_graph(ins, outs, "cnn")
    if opt:
        G.optimise(g)
    assign(ins, values(dt, 10, B=4))
    G.capture_graph(g)
    for rep in range(4):
        vals = values(dt, 20 + rep, B=4)
        assign(ins, vals)
        got = [t.clone() for t in G.replay_graph(g)]
        ref = eager(vals)              # eager work between replays
        torch.cuda.synchronize()
        for a, b in zip(got, ref):
            assert torch.equal(a, b), rep


def eager_fused(vals):
    # the optimised graph's operators, one by one: conv + bias + relu in the
    # fused kernel (bias widened to FP32), then the same pooling and passes
    x, w1, b1, w2, dy = vals
    acc = torch.float64 if x.dtype == torch.float64 else torch.float32
    h = nd.conv2d_bias_act(x, w1, b1.to(acc), [1, 1], SAME, act="relu")
    p = nd.max_pool2d(h, [2, 2], [2, 2])
    y = nd.avg_pool2d(nd.conv2d(p, w2, [1, 1]), [2, 2], [2, 2], VALID)
    dw = nd.conv2d_backward_kernel(p, w2, [1, 1], dy)
    dp = nd.conv2d_backward_input(p, w2, [1, 1], dy)
    dh = nd.max_pool2d_backward(SAME, h, [2, 2], [2, 2], dp)
    up = nd.upsampling2d(p, [2, 2])
    return [y, dw, dh, up]


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [torch.bfloat16, torch.float16])
def test_16bit_graphs_equal_eager(dt):
    # 16-bit: the graph as built equals the unfused operators (torch's 16-bit
    # add and fmax), the optimised graph the fused kernel; both bit for bit
    for opt, ref_fn in ((False, eager), (True, eager_fused)):
        ins, outs = build(dt)
        g = G.make_graph(ins, outs, "cnn16")
        if opt:
            G.optimise(g)
        vals = values(dt, 3)
        assign(ins, vals)
        for a, b in zip(G.eval_graph(g), ref_fn(vals)):
            assert torch.equal(a, b), opt
    # and the fused 16-bit layer is close to the FP32 one before any max-pool
    # routing (a one-ulp change in h can move a max to another cell)
    x, w1, b1 = values(dt, 3)[:3]
    h16 = nd.conv2d_bias_act(x, w1, b1.float(), [1, 1], SAME, act="relu").double()
    h32 = nd.conv2d_bias_act(x.float(), w1.float(), b1.float(), [1, 1], SAME, act="relu").double()
    assert (h16 - h32).norm() / h32.norm() < 1e-2


@pytest.mark.gpu
def test_in_place_variants_equal_allocating_calls():
    rng = np.random.default_rng(4)
    x = rng.standard_normal([2, 9, 11, 8]).astype(np.float32)
    w = (rng.standard_normal([3, 2, 8, 6]) / 5).astype(np.float32)
    y = nd.conv2d(x, w, [2, 1])
    out = np.empty_like(y)
    assert nd.conv2d_(out, x, w, [2, 1]) is None and np.array_equal(out, y)
    dy = rng.standard_normal(y.shape).astype(np.float32)
    for f, f_, args in ((nd.conv2d_backward_input, nd.conv2d_backward_input_, (x, w, [2, 1], dy)),
                        (nd.conv2d_backward_kernel, nd.conv2d_backward_kernel_, (x, w, [2, 1], dy))):
        ref = f(*args)
        o = np.empty_like(ref)
        f_(o, *args)
        assert np.array_equal(o, ref)
    xt, wt = torch.from_numpy(x).to(DEV), torch.from_numpy(w).to(DEV)
    o = torch.empty(y.shape, device=DEV)
    nd.conv2d_(o, xt, wt, [2, 1])
    assert np.array_equal(o.cpu().numpy(), y)
    p = nd.max_pool2d(xt, [2, 3], [2, 2])
    o = torch.empty_like(p)
    nd.max_pool2d_(o, xt, [2, 3], [2, 2])
    assert torch.equal(o, p)
    x1 = x[:, 0]
    y1 = nd.conv1d(x1, w[0], [1])
    o1 = np.empty_like(y1)
    nd.conv1d_(o1, x1, w[0], [1])
    assert np.array_equal(o1, y1)


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [torch.bfloat16, torch.float16])
def test_pooling_16bit_rounds_the_fp32_result_once(dt):
    # 16-bit pooling compares and adds in FP32 in Owl's order and rounds once,
    # so it equals the FP32 kernels on the widened inputs, rounded
    g = torch.Generator(device=DEV)
    g.manual_seed(8)
    x = torch.randn(2, 13, 11, 24, device=DEV, generator=g).to(dt)
    x[0, 3, 4, :5] = float("-inf")
    for kern, stride, pad in (([2, 2], [2, 2], VALID), ([3, 3], [2, 1], SAME), ([3, 2], [1, 2], SAME)):
        for f in (nd.max_pool2d, nd.avg_pool2d):
            y = f(x, kern, stride, pad)
            ref = f(x.float(), kern, stride, pad).to(dt)
            assert torch.equal(y, ref), (f.__name__, kern)
        y = nd.max_pool2d(x, kern, stride, pad)
        dy = torch.randn(y.shape, device=DEV, generator=g).to(dt)
        for f in (nd.max_pool2d_backward, nd.avg_pool2d_backward):
            dx = f(pad, x, kern, stride, dy)
            ref = f(pad, x.float(), kern, stride, dy.float()).to(dt)
            assert torch.equal(dx, ref), (f.__name__, kern)
        ya, arg = nd.max_pool2d_argmax(x, kern, stride, pad)
        ra, rarg = nd.max_pool2d_argmax(x.float(), kern, stride, pad)
        assert torch.equal(ya, ra.to(dt)) and torch.equal(arg, rarg)


@pytest.mark.gpu
def test_captured_fused_pool_survives_larger_library_calls():
    """An FP32 conv + bias + relu + 2x2 max-pool layer without y runs the
    convolution into a library scratch buffer and then the pooling kernel, so
    a captured graph of it holds that buffer's address.  Larger calls of the
    same kind made after the capture need a larger buffer; the captured one
    must stay valid: every replay equals the eager layer bit for bit."""
    dt = torch.float32
    gx, gw, gb = G.var_arr("x", [2, 16, 16, 32], dt), G.var_arr("w", [3, 3, 32, 64], dt), G.var_arr("b", [64], dt)
    gy = G.max_pool2d(G.relu(G.add(G.conv2d(gx, gw, [1, 1]), gb)), [2, 2], [2, 2], VALID)
    g = G.optimise(G.make_graph([gx, gw, gb], [gy], "layer"))
    assert "conv2d_bias_act_maxpool" in [n.op for n in g.order]
    assign([gx, gw, gb], values(dt, 40)[:3])
    G.capture_graph(g)
    gen = torch.Generator(device=DEV)
    gen.manual_seed(41)
    for rep in range(3):
        # a larger keep_y=False layer between replays, growing each time
        big = torch.randn(4 + 4 * rep, 32, 32, 32, device=DEV, generator=gen)
        wv, bv = values(dt, 42)[1:3]
        _, pb = nd.conv2d_bias_act_maxpool(big, wv, bv, [1, 1], SAME, keep_y=False)
        ref_big = nd.max_pool2d(nd.conv2d_bias_act(big, wv, bv, [1, 1], SAME), [2, 2], [2, 2], VALID)
        vals = values(dt, 50 + rep)[:3]
        assign([gx, gw, gb], vals)
        got = G.replay_graph(g)[0].clone()
        ref = nd.max_pool2d(nd.conv2d_bias_act(vals[0], vals[1], vals[2], [1, 1], SAME), [2, 2], [2, 2], VALID)
        torch.cuda.synchronize()
        assert torch.equal(pb, ref_big), rep
        assert torch.equal(got, ref), rep


def _block(dt, with_h2):
    x, w1, b1 = G.var_arr("x", [2, 14, 14, 16], dt), G.var_arr("w1", [3, 3, 16, 32], dt), G.var_arr("b1", [32], dt)
    w2, b2 = G.var_arr("w2", [3, 3, 32, 32], dt), G.var_arr("b2", [32], dt)
    h1 = G.relu(G.add(G.conv2d(x, w1, [1, 1]), b1))
    h2 = G.relu(G.add(G.conv2d(h1, w2, [1, 1]), b2))
    p = G.max_pool2d(h2, [2, 2], [2, 2], VALID)
    return [x, w1, b1, w2, b2], ([p, h2] if with_h2 else [p])


@pytest.mark.gpu
@pytest.mark.parametrize("dt,tol", [(torch.float32, 1e-5), (torch.float64, 1e-12)], ids=["fp32", "fp64"])
@pytest.mark.parametrize("opt", [False, True], ids=["built", "optimised"])
def test_graph_block_matches_oracle(dt, tol, opt):
    """A VGG block (conv + bias + relu twice, then a 2x2 stride-2 VALID
    max-pool) evaluated as a graph, as built and after ``optimise`` (which
    fuses the layers and, when h2 is not an output, the pooling), against the
    FP64 oracle's composition (owl_computation_cpu_eval.ml:180-190 dispatches
    the same operators to the CPU kernels): the pooled output within tol of
    the oracle; and, with h2 kept as an output, the pooled output equal bit
    for bit to the oracle's max-pool of the GPU's h2."""
    import oracle_ndarray as on
    g = torch.Generator(device=DEV)
    g.manual_seed(60)
    r = lambda *s, scale=1.0: (torch.randn(*s, device=DEV, generator=g, dtype=torch.float64) * scale).to(dt)
    vals = [r(2, 14, 14, 16), r(3, 3, 16, 32, scale=0.12), r(32, scale=0.2), r(3, 3, 32, 32, scale=0.08),
            r(32, scale=0.2)]
    h = [v.double().cpu().numpy() for v in vals]
    r1 = np.maximum(on.conv2d(h[0], h[1], [1, 1], SAME) + h[2], 0.0)
    r2 = np.maximum(on.conv2d(r1, h[3], [1, 1], SAME) + h[4], 0.0)
    rp = on.max_pool2d(r2, [2, 2], [2, 2], VALID)
    for with_h2 in (False, True):
        ins, outs = _block(dt, with_h2)
        graph = G.make_graph(ins, outs, "block")
        if opt:
            G.optimise(graph)
            ops = [n.op for n in graph.order]
            assert ops.count("conv2d_bias_act") == (2 if with_h2 else 1), ops
            assert ("conv2d_bias_act_maxpool" in ops) == (not with_h2), ops
        assign(ins, vals)
        got = G.eval_graph(graph)
        torch.cuda.synchronize()
        p = got[0].cpu().numpy()
        assert p.dtype == np.dtype(str(dt).split(".")[1])
        assert float(np.linalg.norm(p - rp) / np.linalg.norm(rp)) < tol, with_h2
        if with_h2:
            h2 = got[1].cpu().numpy()
            assert float(np.linalg.norm(h2 - r2) / np.linalg.norm(r2)) < tol
            assert np.array_equal(p, on.max_pool2d(h2, [2, 2], [2, 2], VALID))
