"""The GPU computation-graph engine's host side, without a GPU: shape
inference (against the ndarray wrappers' own geometry), evaluation order, the
block plan, the conv + bias + relu fusion, and the checks on ``out`` arrays
of the in-place ndarray variants.

Plan safety is checked by brute force over every pair of nodes: two results
share a block only if their live ranges (from the node that makes one to its
last consumer; graph outputs and variables to the end) do not overlap, the
one exception being relu / add written over a first parent that dies at them."""
import itertools

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from owl_amd import cgraph as G
from owl_amd import ndarray as nd
from owl_amd.shapes import SAME, VALID, InvalidArgument


def live_ranges(g):
    idx = {n.id: i for i, n in enumerate(g.order)}
    last = {}
    for i, n in enumerate(g.order):
        for p in n.parents:
            last[p.id] = i
    end = len(g.order)
    keep = {n.id for n in g.outputs} | {n.id for n in g.order if n.op == "var"}
    # variables are written before the run starts: live from the beginning
    return {n.id: (-1 if n.op == "var" else idx[n.id], end if n.id in keep else last.get(n.id, idx[n.id]))
            for n in g.order if n.op != "const"}


def check_plan(g):
    rng = live_ranges(g)
    byid = {n.id: n for n in g.order}
    for n in g.order:
        if n.op == "const":
            continue
        assert g.block_bytes[g.plan[n.id]] >= n.nbytes()
    for a, b in itertools.combinations([n for n in g.order if n.op != "const"], 2):
        if g.plan[a.id] != g.plan[b.id]:
            continue
        (a0, a1), (b0, b1) = rng[a.id], rng[b.id]
        if a1 < b0 or b1 < a0:
            continue
        # the only sanctioned overlap: b overwrites its first parent a, which dies at b
        assert b.op in ("relu", "add") and b.parents[0] is a and a1 == b0, (a, b)
        assert all(q is not a for q in b.parents[1:])
    for n in g.order:  # parents come before their consumers
        for p in n.parents:
            assert g.order.index(p) < g.order.index(n)
    return byid


def small_cnn(dt=torch.float32, B=2):
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


@pytest.mark.parametrize("seed", range(6))
def test_shapes_match_ndarray_geometry(seed):
    rng = np.random.default_rng(seed)
    H, W = int(rng.integers(5, 30)), int(rng.integers(5, 30))
    kh, kw = int(rng.integers(1, 5)), int(rng.integers(1, 5))
    s = [int(rng.integers(1, 4)), int(rng.integers(1, 4))]
    r = [int(rng.integers(1, 3)), int(rng.integers(1, 3))]
    pad = [SAME, VALID][seed % 2]
    if (H - (kh - 1) * r[0] < 1 or W - (kw - 1) * r[1] < 1) and pad == VALID:
        pad = SAME
    x = G.var_arr("x", [3, H, W, 8])
    w = G.var_arr("w", [kh, kw, 8, 5])
    geo = nd.conv_problem(x.shape, w.shape, s, r, pad)
    assert G.dilated_conv2d(x, w, s, r, pad).shape == [3, *geo.sp_out, 5]
    geo1 = nd.conv_problem(x.shape, w.shape, s, None, pad)
    assert G.conv2d(x, w, s, pad).shape == [3, *geo1.sp_out, 5]
    ps = G.max_pool2d(x, [kh, kw], s, SAME).shape
    assert ps == [3, -(-H // s[0]), -(-W // s[1]), 8]
    assert G.upsampling2d(x, [2, 3]).shape == [3, 2 * H, 3 * W, 8]


def test_plan_is_safe_and_reuses_blocks():
    ins, outs = small_cnn()
    g = G.make_graph(ins, outs, "cnn")
    byid = check_plan(g)
    ops = [n.op for n in g.order]
    # conv -> add -> relu run in one block (add and relu overwrite in place)
    conv1 = next(n for n in g.order if n.op == "conv2d" and n.parents[0].name == "x")
    add = next(n for n in g.order if n.op == "add")
    rel = next(n for n in g.order if n.op == "relu")
    assert g.plan[conv1.id] == g.plan[add.id] == g.plan[rel.id]
    # fewer blocks than results, and fewer bytes than one block each
    results = [n for n in g.order if n.op != "const"]
    assert len(g.block_bytes) < len(results)
    assert sum(g.block_bytes) < sum(n.nbytes() for n in results)
    assert ops.index("max_pool2d_backward") > ops.index("conv2d_backward_input")
    assert byid


def test_missing_variable_is_an_error():
    ins, outs = small_cnn()
    with pytest.raises(InvalidArgument):
        G.make_graph(ins[:-1], outs, "cnn")


def test_optimise_fuses_conv_bias_relu():
    ins, outs = small_cnn()
    g = G.optimise(G.make_graph(ins, outs, "cnn"))
    ops = [n.op for n in g.order]
    assert ops.count("conv2d_bias_act") == 1 and "add" not in ops and "relu" not in ops
    fused = next(n for n in g.order if n.op == "conv2d_bias_act")
    assert fused.params["act"] == "relu" and [p.name for p in fused.parents] == ["x", "w1", "b1"]
    check_plan(g)
    # the max-pool backward still reads the fused node's result as h
    mpb = next(n for n in g.order if n.op == "max_pool2d_backward")
    assert mpb.parents[0] is fused


def test_optimise_respects_other_consumers_and_outputs():
    x = G.var_arr("x", [1, 8, 8, 16])
    w = G.var_arr("w", [3, 3, 16, 16])
    b = G.var_arr("b", [16])
    c = G.conv2d(x, w, [1, 1])
    a = G.add(c, b)
    r = G.relu(a)
    # c is also an output: nothing fuses
    g = G.optimise(G.make_graph([x, w, b], [r, c], "g"))
    assert "conv2d_bias_act" not in [n.op for n in g.order]
    # a is also consumed elsewhere: conv + add fuse (act none), relu stays
    x2 = G.var_arr("x", [1, 8, 8, 16])
    c2 = G.conv2d(x2, w, [1, 1])
    a2 = G.add(c2, b)
    r2 = G.relu(a2)
    m2 = G.max_pool2d(a2, [2, 2], [2, 2])
    g2 = G.optimise(G.make_graph([x2, w, b], [r2, m2], "g2"))
    ops = [n.op for n in g2.order]
    assert ops.count("conv2d_bias_act") == 1 and "relu" in ops
    assert next(n for n in g2.order if n.op == "conv2d_bias_act").params["act"] is None
    # a full-shape add (not a bias) is left alone
    x3 = G.var_arr("x", [1, 8, 8, 16])
    c3 = G.conv2d(x3, w, [1, 1])
    z = G.var_arr("z", c3.shape)
    g3 = G.optimise(G.make_graph([x3, w, z], [G.relu(G.add(c3, z))], "g3"))
    assert "conv2d_bias_act" not in [n.op for n in g3.order]
    for gg in (g, g2, g3):
        check_plan(gg)


def test_operand_checks():
    x = G.var_arr("x", [1, 8, 8, 16])
    with pytest.raises(InvalidArgument):
        G.conv2d(x, G.var_arr("w", [3, 3, 8, 4]), [1, 1])
    with pytest.raises(InvalidArgument):
        G.add(x, G.var_arr("b", [8]))
    with pytest.raises(InvalidArgument):
        G.conv2d(x, G.var_arr("w", [3, 3, 16, 4], torch.float64), [1, 1])


def test_out_argument_checks():
    x = np.zeros([1, 6, 6, 4], np.float32)
    w = np.zeros([3, 3, 4, 5], np.float32)
    with pytest.raises(InvalidArgument):   # wrong shape
        nd.conv2d_(np.zeros([1, 6, 6, 4], np.float32), x, w, [1, 1])
    with pytest.raises(InvalidArgument):   # wrong dtype
        nd.conv2d_(np.zeros([1, 6, 6, 5], np.float64), x, w, [1, 1])
    with pytest.raises(InvalidArgument):   # not contiguous
        nd.conv2d_(np.zeros([1, 6, 6, 10], np.float32)[..., ::2], x, w, [1, 1])
    with pytest.raises(InvalidArgument):   # overlaps the input
        nd.max_pool2d_(x, x, [1, 1], [1, 1])
    with pytest.raises(InvalidArgument):   # 1-D out must be contiguous
        nd.conv1d_(np.zeros([1, 6, 10], np.float32)[..., ::2], x[:, 0], w[0], [1])
    assert nd.conv2d_.__doc__ and "2402" in nd.conv2d_.__doc__


def vgg_block(pool_padding=VALID, extra_out=False):
    x = G.var_arr("x", [2, 8, 8, 16])
    w1, b1 = G.var_arr("w1", [3, 3, 16, 32]), G.var_arr("b1", [32])
    w2, b2 = G.var_arr("w2", [3, 3, 32, 32]), G.var_arr("b2", [32])
    h1 = G.relu(G.add(G.conv2d(x, w1, [1, 1]), b1))
    h2 = G.relu(G.add(G.conv2d(h1, w2, [1, 1]), b2))
    y = G.max_pool2d(h2, [2, 2], [2, 2], pool_padding)
    return G.make_graph([x, w1, b1, w2, b2], [y, h2] if extra_out else [y], "vgg_block")


def test_optimise_fuses_conv_bias_relu_maxpool():
    g = G.optimise(vgg_block())
    ops = [n.op for n in g.order]
    assert ops.count("conv2d_bias_act") == 1 and ops.count("conv2d_bias_act_maxpool") == 1
    assert not {"max_pool2d", "add", "relu", "conv2d"} & set(ops)
    last = next(n for n in g.order if n.op == "conv2d_bias_act_maxpool")
    first = next(n for n in g.order if n.op == "conv2d_bias_act")
    assert last.parents[0] is first and [p.name for p in last.parents[1:]] == ["w2", "b2"]
    assert last.shape == [2, 4, 4, 32] and last.params["act"] == "relu" and g.outputs[0] is last
    check_plan(g)
    # a SAME-padded pool, or a layer output that is also a graph output, stays apart
    for gg in (G.optimise(vgg_block(pool_padding=SAME)), G.optimise(vgg_block(extra_out=True))):
        ops = [n.op for n in gg.order]
        assert "conv2d_bias_act_maxpool" not in ops and ops.count("conv2d_bias_act") == 2 and "max_pool2d" in ops
        check_plan(gg)
