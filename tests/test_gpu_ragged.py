"""Ragged batch (64 seeded conv2d problems of mixed shape, BF16 in, FP32
accumulate) in one persistent launch: equal to the same problems run as
separate launches, stable across hipGraph replays and autotuning (tile edge,
split-K, own launch per shape class), and within 1e-2 of the FP64 oracle on
sampled output pixels; forced split-K plans against the unsplit sums."""
import numpy as np
import pytest

from owl_amd import ndarray as nd
from owl_amd.shapes import c_pad, k_eff

torch = pytest.importorskip("torch")
from owl_amd.group import ConvGroup, ragged_problems  # noqa: E402

DEV = "cuda:0"


def make_inputs(problems, seed):
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    xs, ws = [], []
    for xs_, ks_, *_ in problems:
        xs.append((torch.rand(xs_, device=DEV, generator=g) * 2 - 1).to(torch.bfloat16))
        ws.append((torch.rand(ks_, device=DEV, generator=g) * 2 - 1).to(torch.bfloat16))
    return xs, ws


def sampled_oracle(x, w, stride, rate, pad, y, rng, n=8):
    """FP64 outputs at n random (b, i, j) positions, all output channels."""
    x = x.double().cpu().numpy()
    w = w.double().cpu().numpy()
    B, H, W, Ci = x.shape
    kh, kw, _, Co = w.shape
    Ho, Wo = y.shape[1:3]
    ph = c_pad(stride[0], Ho, k_eff(kh, rate[0]), H) if pad == "SAME" else 0
    pw = c_pad(stride[1], Wo, k_eff(kw, rate[1]), W) if pad == "SAME" else 0
    pts, refs = [], []
    for _ in range(n):
        b, i, j = int(rng.integers(B)), int(rng.integers(Ho)), int(rng.integers(Wo))
        acc = np.zeros(Co)
        for a in range(kh):
            h = i * stride[0] - ph + a * rate[0]
            if not 0 <= h < H:
                continue
            for c in range(kw):
                ww = j * stride[1] - pw + c * rate[1]
                if 0 <= ww < W:
                    acc += x[b, h, ww] @ w[a, c]
        pts.append((b, i, j))
        refs.append(acc)
    return pts, refs


def _rel(a, b):
    a, b = a.double(), b.double()
    return (torch.linalg.norm(a - b) / torch.linalg.norm(b)).item()


@pytest.mark.gpu
def test_ragged_group_matches_separate_launches_and_oracle():
    problems = ragged_problems(seed=4, n=64)
    xs, ws = make_inputs(problems, seed=44)
    group = ConvGroup(problems, torch.bfloat16, torch.bfloat16)
    ys = group.run(xs, ws)
    torch.cuda.synchronize()
    sep = [nd.dilated_conv2d(x, w, p[2], p[3], p[4]) for x, w, p in zip(xs, ws, problems)]
    torch.cuda.synchronize()
    rng = np.random.default_rng(0)
    for (xs_, ks_, st, rt, pad), x, w, y, s in zip(problems, xs, ws, ys, sep):
        assert tuple(y.shape) == tuple(s.shape)
        yd, sd = y.double(), s.double()
        assert (torch.linalg.norm(yd - sd) / torch.linalg.norm(sd)).item() < 1e-2
        pts, refs = sampled_oracle(x, w, st, rt, pad, y, rng)
        for (b, i, j), r in zip(pts, refs):
            got = y[b, i, j].double().cpu().numpy()
            assert np.linalg.norm(got - r) <= 1e-2 * max(np.linalg.norm(r), 1e-6)

    # one graph-captured launch replayed several times, eager work in between
    stream = torch.cuda.Stream()
    ys_g = group.empty_outputs()
    with torch.cuda.stream(stream):
        group.capture(xs, ws, ys_g, stream)
        for rep in range(3):
            for y in ys_g:
                y.zero_()
            group.replay(stream)
            stream.synchronize()
            for a, b in zip(ys_g, ys):
                assert torch.equal(a, b), f"graph replay {rep} differs from the eager grouped launch"
            _ = torch.ones(1024, device=DEV).sum().item()

    # autotuning picks, per shape class, a tile edge (same K order per
    # element: the eager grouped result), split-K (FP32 partials summed in
    # split order: within BF16 rounding of it) or the problem's own launch
    # (the separate launch's result)
    ys_t = group.empty_outputs()
    group.autotune(xs, ws, ys_t)
    edges, _ = group.tile_edges()
    kinds, splits = group.plan()
    assert set(edges) <= {0, 64, 128}
    ys_t = group.run(xs, ws, ys_t)
    torch.cuda.synchronize()
    for i, (a, e, s) in enumerate(zip(ys_t, edges, splits)):
        if s > 1:
            assert e > 0 and _rel(a, ys[i]) < 1e-2, (i, e, s)
        else:
            assert torch.equal(a, sep[i] if e == 0 else ys[i]), (i, e)

    # and a capture of the tuned plan replays the same
    ys_g2 = group.empty_outputs()
    with torch.cuda.stream(stream):
        group.capture(xs, ws, ys_g2, stream)
        for rep in range(3):
            for y in ys_g2:
                y.zero_()
            group.replay(stream)
            stream.synchronize()
            for a, b in zip(ys_g2, ys_t):
                assert torch.equal(a, b), f"tuned graph replay {rep} differs from the tuned run"
            _ = torch.ones(1024, device=DEV).sum().item()


@pytest.mark.gpu
def test_group_own_launch_classes():
    """Large stride-1 3x3 layers beside small ones: whatever the tuner picks,
    every problem equals its separate launch's result when it has its own
    launch, the eager grouped result when it stays unsplit in the persistent
    launch, and is within 1e-2 of the separate launch in any case."""
    problems = [([4, 56, 56, 128], [3, 3, 128, 256], [1, 1], [1, 1], "SAME"),
                ([2, 56, 56, 128], [3, 3, 128, 256], [1, 1], [1, 1], "SAME"),
                ([3, 14, 14, 32], [1, 1, 32, 48], [2, 2], [1, 1], "VALID"),
                ([1, 7, 7, 16], [3, 3, 16, 24], [1, 1], [2, 2], "SAME")]
    xs, ws = make_inputs(problems, seed=45)
    group = ConvGroup(problems, torch.bfloat16, torch.bfloat16)
    ys = group.run(xs, ws)
    sep = [nd.dilated_conv2d(x, w, p[2], p[3], p[4]) for x, w, p in zip(xs, ws, problems)]
    ys_t = group.empty_outputs()
    group.autotune(xs, ws, ys_t)
    edges, total = group.tile_edges()
    _, splits = group.plan()
    ys_t = group.run(xs, ws, ys_t)
    torch.cuda.synchronize()
    assert total >= 0
    for i, (a, e, s) in enumerate(zip(ys_t, edges, splits)):
        if s == 1:
            assert torch.equal(a, sep[i] if e == 0 else ys[i]), (i, e)
        assert _rel(a, sep[i]) < 1e-2


@pytest.mark.gpu
def test_group_split_k_plan():
    """Problems forced to split-K in the persistent launch: with FP32 outputs
    the split sums (FP32 partials, summed in split order) equal the unsplit
    grouped sums to FP32 rounding, BF16 outputs stay within 1e-2 of the FP64
    oracle, unsplit problems keep their bits, and a captured run replays the
    split plan exactly."""
    problems = [([1, 7, 7, 512], [3, 3, 512, 256], [1, 1], [1, 1], "SAME"),     # K 4608: 144 K-steps
                ([2, 14, 14, 128], [5, 5, 128, 64], [2, 2], [1, 1], "SAME"),    # K 3200
                ([3, 9, 9, 20], [3, 3, 20, 24], [1, 1], [2, 2], "VALID"),       # unaligned channels, K 180
                ([2, 28, 28, 64], [1, 1, 64, 32], [1, 1], [1, 1], "SAME")]      # K 64
    xs, ws = make_inputs(problems, seed=46)
    for out_dt in (torch.float32, torch.bfloat16):
        group = ConvGroup(problems, torch.bfloat16, out_dt)
        group.set_plan([0, 1, 1, 0], [1, 1, 1, 1])
        base = group.run(xs, ws)
        group.set_plan([0, 1, 1, 0], [16, 4, 3, 1])
        kinds, splits = group.plan()
        assert kinds == [0, 1, 1, 0]
        # 144 K-steps in 16 splits of 9; 100 in 4 of 25; K 180 (5.6 steps) in 3 of 2 steps; unsplit
        assert splits == [16, 4, 3, 1], splits
        ys = group.run(xs, ws)
        torch.cuda.synchronize()
        assert torch.equal(ys[3], base[3])
        rng = np.random.default_rng(1)
        for i in range(3):
            if out_dt == torch.float32:
                assert _rel(ys[i], base[i]) < 1e-6, i
            xs_, ks_, st, rt, pad = problems[i]
            pts, refs = sampled_oracle(xs[i], ws[i], st, rt, pad, ys[i], rng)
            for (b, r_, c), r in zip(pts, refs):
                got = ys[i][b, r_, c].double().cpu().numpy()
                assert np.linalg.norm(got - r) <= 1e-2 * max(np.linalg.norm(r), 1e-6), (i, out_dt)
        stream = torch.cuda.Stream()
        ys_g = group.empty_outputs()
        with torch.cuda.stream(stream):
            group.capture(xs, ws, ys_g, stream)
            for rep in range(3):
                for y in ys_g:
                    y.zero_()
                group.replay(stream)
                stream.synchronize()
                for a, b in zip(ys_g, ys):
                    assert torch.equal(a, b), f"split-K graph replay {rep} differs"
                _ = torch.ones(1024, device=DEV).sum().item()


@pytest.mark.gpu
def test_group_plan_arguments():
    problems = [([1, 8, 8, 16], [3, 3, 16, 16], [1, 1], [1, 1], "SAME")]
    group = ConvGroup(problems, torch.bfloat16, torch.bfloat16)
    for kinds, splits in (([3], [1]), ([0], [0]), ([0], [17]), ([2], [2])):
        with pytest.raises(Exception):
            group.set_plan(kinds, splits)
    group.set_plan([2], [1])
    assert group.plan() == ([2], [1])
    assert group.tile_edges() == ([0], 0)
