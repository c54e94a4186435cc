#!/usr/bin/env python3
"""Benchmark: Owl's conv2d layer forward + backward_input + backward_kernel on
MI355X, BF16 MFMA path against the plain FP32 path.

Workload (the VGG-style layer): x [64,56,56,64] ~ N(0,1), kernel [3,3,64,128]
He-normal (std sqrt(2/576)), stride 1, SAME; output gradient [64,56,56,128]
~ N(0,1); then max_pool2d 2x2 stride 2 VALID forward (with argmax) and
backward on the FP32 layer output.  A "step" is the three conv passes
(2 * 3 * 29.595 GFLOP = 88.785 GFLOP) over resident HBM buffers.

Prints one JSON line (rank 0, N = 1).  ``value`` is the BF16 step's TFLOP/s.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B, H, W, CI, CO, K = 64, 56, 56, 64, 128, 3
FLOP_PASS = 2.0 * B * H * W * CO * K * K * CI  # 29.595 GFLOP
PEAK = {"bf16": 2.5e15, "f16": 2.5e15, "f32": 157.3e12}
HBM_PEAK = 8.0e12
SPINUP_S = 0.25  # untimed work before each timed region, after the W warmup steps
# HBM bytes per launch of the BF16 step's kernels, from two rocprofv3 --pmc
# runs (FETCH_SIZE, then WRITE_SIZE) scaled by what the same counters read
# for a plain copy of known size (x2.0 and x1.0); the runs and their command
# are in profiles/conv2d_vgg_layer_bf16_hbm_traffic.md
TRAFFIC = {
    ("conv_halo_kernel", "fwd"): 13253.5 * 2048 + 50209.2 * 1024,
    ("conv_halo_kernel", "bwd_input"): 26504.1 * 2048 + 27330.2 * 1024,
    ("conv_wgrad_kernel", "bwd_kernel"): 112607.0 * 2048 + 24509.8 * 1024,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 oracle error check")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the benchmark measures one GPU; --gpus {args.gpus} is not supported "
              "(the layer would shard over the batch, with one all-reduce of dw; that is not implemented)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from owl_amd import _lib
    from owl_amd import ndarray as nd
    from owl_amd.shapes import SAME, VALID

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    x32 = torch.randn(B, H, W, CI, device=dev, generator=g)
    w32 = torch.randn(K, K, CI, CO, device=dev, generator=g) * (2.0 / (K * K * CI)) ** 0.5
    dy32 = torch.randn(B, H, W, CO, device=dev, generator=g)
    xb, wb, dyb = x32.bfloat16(), w32.bfloat16(), dy32.bfloat16()

    def step(x, w, dy, out_w):
        y = nd.conv2d(x, w, [1, 1], SAME)
        dx = nd.conv2d_backward_input(x, w, [1, 1], dy)
        dw = nd.conv2d_backward_kernel(x, w, [1, 1], dy, out_dtype=out_w)
        return y, dx, dw

    def timed(x, w, dy, out_w, steps, warmup, kernel_times=False):
        """Wall-clock K steps (the measured value), then -- outside that
        region -- each pass K times back to back between two HIP events,
        and (BF16) K more steps with the library's per-kernel timing (the
        events stamped with each kernel's own start and end)."""
        # W warmup steps, then untimed steps until SPINUP_S has passed: the
        # step runs ~9% slower until the GPU has been busy for ~10 ms
        w0 = time.perf_counter()
        for _ in range(warmup):
            step(x, w, dy, out_w)
        torch.cuda.synchronize()
        while time.perf_counter() - w0 < SPINUP_S:
            for _ in range(8):
                step(x, w, dy, out_w)
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(x, w, dy, out_w)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # per pass: K back-to-back calls between events on the launch stream
        passes = (lambda: nd.conv2d(x, w, [1, 1], SAME),
                  lambda: nd.conv2d_backward_input(x, w, [1, 1], dy),
                  lambda: nd.conv2d_backward_kernel(x, w, [1, 1], dy, out_dtype=out_w))
        per = []
        for f in passes:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(steps):
                f()
            e1.record()
            torch.cuda.synchronize()
            per.append(e0.elapsed_time(e1) / steps)
        kt = None
        if kernel_times:
            _lib.kernel_timing(True)
            for _ in range(steps):
                step(x, w, dy, out_w)
            torch.cuda.synchronize()
            kt = {}
            for cls, tag in ((_lib.KT_FWD, "fwd"), (_lib.KT_BWD_INPUT, "bwd_input"),
                             (_lib.KT_BWD_KERNEL, "bwd_kernel"), (_lib.KT_AUX, "aux")):
                ms, n, name = _lib.kernel_time(cls)
                kt[tag] = {"kernel": name, "launches": n, "avg_ms": ms / max(n, 1)}
            _lib.kernel_timing(False)
        return wall, per, kt

    wall_b, per_b, kt_b = timed(xb, wb, dyb, torch.float32, args.steps, args.warmup, kernel_times=True)
    wall_f, per_f, _ = timed(x32, w32, dy32, torch.float32, args.steps, args.warmup)
    tf_b = 3 * FLOP_PASS * args.steps / wall_b / 1e12
    tf_f = 3 * FLOP_PASS * args.steps / wall_f / 1e12

    # the conv layer with its bias and relu (Owl's Conv2D + relu): fused into
    # the BF16 forward kernel's epilogue, against conv then separate ops
    bias = torch.randn(CO, device=dev, generator=g) * 0.1
    zero = torch.zeros((), device=dev, dtype=torch.bfloat16)

    def t_loop(f, reps=20):
        for _ in range(3):
            f()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3
    fused = {
        "fused_us": round(t_loop(lambda: nd.conv2d_bias_act(xb, wb, bias, [1, 1], SAME)), 1),
        "fused_keep_pre_us": round(t_loop(lambda: nd.conv2d_bias_act(xb, wb, bias, [1, 1], SAME,
                                                                     keep_pre=True)), 1),
        "unfused_us": round(t_loop(lambda: torch.fmax(nd.conv2d(xb, wb, [1, 1], SAME) + bias.bfloat16(), zero)),
                            1),
        "layer": "BF16 relu(conv2d(x, w) + b), x[64,56,56,64] w[3,3,64,128]",
    }
    # the same layer followed by the 2x2 stride-2 max-pool: pooled in the
    # forward kernel's epilogue (with y, and without it for inference)
    # against the fused layer then the pooling kernel
    hb = nd.conv2d_bias_act(xb, wb, bias, [1, 1], SAME)
    fused_pool = {
        "fused_us": round(t_loop(lambda: nd.conv2d_bias_act_maxpool(xb, wb, bias, [1, 1], SAME)), 1),
        "fused_no_y_us": round(t_loop(lambda: nd.conv2d_bias_act_maxpool(xb, wb, bias, [1, 1], SAME,
                                                                          keep_y=False)), 1),
        "unfused_us": round(t_loop(lambda: nd.max_pool2d(nd.conv2d_bias_act(xb, wb, bias, [1, 1], SAME), [2, 2],
                                                         [2, 2], VALID)), 1),
        "pool_alone_us": round(t_loop(lambda: nd.max_pool2d(hb, [2, 2], [2, 2], VALID)), 1),
    }
    # its backward up to dz and db: the max-pool backward and the relu / bias
    # adjoint in one pass against the two calls
    hy, hp, hpre = nd.conv2d_bias_act_maxpool(xb, wb, bias, [1, 1], SAME, keep_pre=True)
    dpool = torch.randn(hp.shape, device=dev, generator=g).bfloat16()
    fused_pool["bwd_fused_us"] = round(t_loop(lambda: nd.maxpool_bias_act_backward(hy, hpre, dpool)), 1)
    fused_pool["bwd_unfused_us"] = round(t_loop(lambda: nd.bias_act_backward(
        hpre, nd.max_pool2d_backward(VALID, hy, [2, 2], [2, 2], dpool), act="relu")), 1)
    fused_pool["layer"] = "BF16 max_pool2x2(relu(conv2d(x, w) + b)), x[64,56,56,64] w[3,3,64,128]"
    del hb, hy, hp, hpre, dpool

    # the FP8 e4m3 forward of the same layer (per-tensor amax scales, FP8
    # operands resident, BF16 output), quantisation outside the timed loop
    import ctypes
    xq = torch.empty(xb.numel(), dtype=torch.uint8, device=dev)
    wq = torch.empty(wb.numel(), dtype=torch.uint8, device=dev)
    sx = torch.empty(1, dtype=torch.float32, device=dev)
    sw = torch.empty(1, dtype=torch.float32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for t, q, sc in ((xb, xq, sx), (wb, wq, sw)):
        _lib.check(_lib.lib().owl_amd_fp8_quantize(_lib.BF16, ctypes.c_void_p(t.data_ptr()), t.numel(),
                                                   ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(sc.data_ptr()),
                                                   stream))
    y8 = torch.empty(B, H, W, CO, dtype=torch.bfloat16, device=dev)
    g8 = nd._Conv(B, [H, W], CI, [K, K], CO, [H, W], [1, 1], [1, 1], [1, 1]).desc()
    fp8_call = lambda: _lib.check(_lib.lib().owl_amd_conv_fwd_fp8(
        ctypes.byref(g8), _lib.BF16, ctypes.c_void_p(xq.data_ptr()), ctypes.c_void_p(wq.data_ptr()),
        ctypes.c_void_p(sx.data_ptr()), ctypes.c_void_p(sw.data_ptr()), ctypes.c_void_p(y8.data_ptr()), stream))
    t8 = t_loop(fp8_call)
    y8_rel = float(torch.linalg.vector_norm(y8[:2].double() - nd.conv2d(x32[:2].contiguous(), w32, [1, 1], SAME)
                                            .double()) / torch.linalg.vector_norm(
        nd.conv2d(x32[:2].contiguous(), w32, [1, 1], SAME).double()))
    fp8 = {"fwd_us": round(t8, 1), "tflops": round(FLOP_PASS / (t8 * 1e-6) / 1e12, 1),
           "frac_of_fp8_peak": round(FLOP_PASS / (t8 * 1e-6) / 5.0e15, 4),
           "rel_l2_vs_fp32_conv_2_images": round(y8_rel, 4),
           "what": "FP8 e4m3 forward of the VGG layer (halo kernel, 32x32x64 f8f6f4 MFMA, kernel repacked to "
                   "[tap][co][ci] per call, included), quantised operands resident, BF16 output"}
    del xq, wq, y8

    # a VGG block as a computation graph (BF16): conv + bias + relu twice,
    # then a 2x2 max-pool; eager unfused ndarray calls, the optimised graph
    # (both layers fused) evaluated eagerly, and its HIP-graph replay
    from owl_amd import cgraph as G
    w2b = (torch.randn(K, K, CO, CO, device=dev, generator=g) * (2.0 / (K * K * CO)) ** 0.5).bfloat16()
    b2 = (torch.randn(CO, device=dev, generator=g) * 0.1).bfloat16()
    b1 = bias.bfloat16()
    gx, gw1, gb1 = G.var_arr("x", [B, H, W, CI], torch.bfloat16), G.var_arr("w1", [K, K, CI, CO], torch.bfloat16), \
        G.var_arr("b1", [CO], torch.bfloat16)
    gw2, gb2 = G.var_arr("w2", [K, K, CO, CO], torch.bfloat16), G.var_arr("b2", [CO], torch.bfloat16)
    h1 = G.relu(G.add(G.conv2d(gx, gw1, [1, 1]), gb1))
    gy = G.max_pool2d(G.relu(G.add(G.conv2d(h1, gw2, [1, 1]), gb2)), [2, 2], [2, 2], VALID)
    graph = G.optimise(G.make_graph([gx, gw1, gb1, gw2, gb2], [gy], "vgg_block"))
    for n_, v_ in zip([gx, gw1, gb1, gw2, gb2], [xb, wb, b1, w2b, b2]):
        G.assign_arr(n_, v_)

    def block_eager():
        h = torch.fmax(nd.conv2d(xb, wb, [1, 1], SAME) + b1, zero)
        h = torch.fmax(nd.conv2d(h, w2b, [1, 1], SAME) + b2, zero)
        return nd.max_pool2d(h, [2, 2], [2, 2], VALID)
    t_eager = t_loop(block_eager)
    t_graph = t_loop(lambda: G.eval_graph(graph))
    G.capture_graph(graph)
    t_replay = t_loop(lambda: G.replay_graph(graph))
    gflop_block = (FLOP_PASS + 2.0 * B * H * W * CO * K * K * CO) / 1e9
    cgraph_block = {
        "eager_unfused_us": round(t_eager, 1), "graph_eval_us": round(t_graph, 1),
        "graph_replay_us": round(t_replay, 1),
        "graph_replay_tflops": round(gflop_block * 1e9 / (t_replay * 1e-6) / 1e12, 1),
        "graph_blocks_mb": round(graph.planned_bytes() / 1e6, 1),
        "block": "BF16 max_pool2x2(relu(conv(relu(conv(x, w1) + b1), w2) + b2)), x[64,56,56,64] "
                 "w1[3,3,64,128] w2[3,3,128,128]; replay includes copying the 5 inputs into the graph",
    }

    # pooling on the FP32 layer output
    y32 = nd.conv2d(x32, w32, [1, 1], SAME)
    gp = torch.randn(B, H // 2, W // 2, CO, device=dev, generator=g)
    for _ in range(3):
        nd.max_pool2d_argmax(y32, [2, 2], [2, 2], VALID)
        nd.max_pool2d_backward(VALID, y32, [2, 2], [2, 2], gp)
    torch.cuda.synchronize()
    pe = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    reps = 10
    pe[0].record()
    for _ in range(reps):
        nd.max_pool2d(y32, [2, 2], [2, 2], VALID)
    pe[1].record()
    for _ in range(reps):
        nd.max_pool2d_argmax(y32, [2, 2], [2, 2], VALID)
    pe[2].record()
    for _ in range(reps):
        nd.max_pool2d_backward(VALID, y32, [2, 2], [2, 2], gp)
    pe[3].record()
    torch.cuda.synchronize()
    ny, no = B * H * W * CO * 4, B * (H // 2) * (W // 2) * CO * 4
    t_pf, t_pa, t_pb = (pe[i].elapsed_time(pe[i + 1]) / reps * 1e-3 for i in range(3))
    pool = {"fwd_GBps": round((ny + no) / t_pf / 1e9, 1),
            "argmax_GBps": round((ny + no + 2 * no) / t_pa / 1e9, 1),
            "bwd_GBps": round((ny + no + ny) / t_pb / 1e9, 1),
            "fwd_frac_hbm": round((ny + no) / t_pf / HBM_PEAK, 3),
            "bwd_frac_hbm": round((ny + no + ny) / t_pb / HBM_PEAK, 3)}

    # errors of the timed full-size calls against FP64 on the unrounded FP32
    # inputs: y and dx on the first 2 images against torch's FP64
    # convolution and its input gradient (NCHW views of the NHWC tensors; SAME
    # for a 3x3 stride-1 kernel pads 1 on each side); dw on 256 sampled
    # entries, each an FP64 dot product over all 200704 pixels
    errs = None
    if not args.no_check:
        import torch.nn.functional as F
        n = 2
        x64 = x32[:n].double().permute(0, 3, 1, 2).contiguous()
        w64 = w32.double().permute(3, 2, 0, 1).contiguous()          # [co, ci, kh, kw]
        g64 = dy32[:n].double().permute(0, 3, 1, 2).contiguous()
        ry = F.conv2d(x64, w64, padding=1).permute(0, 2, 3, 1)
        rdx = torch.nn.grad.conv2d_input(x64.shape, w64, g64, padding=1).permute(0, 2, 3, 1)
        rel = lambda a, b: float(torch.linalg.vector_norm(a.double() - b) / torch.linalg.vector_norm(b))
        gi = torch.Generator()
        gi.manual_seed(99)
        flat = torch.randperm(K * K * CI * CO, generator=gi)[:256]
        idx = torch.stack([flat // (K * CI * CO), flat // (CI * CO) % K, flat // CO % CI, flat % CO], 1).tolist()
        xpad = F.pad(x32.double(), (0, 0, 1, 1, 1, 1))
        dyd = dy32.double()
        rdw = torch.stack([(xpad[:, a:a + H, c:c + W, q] * dyd[..., o]).sum() for a, c, q, o in idx])
        del xpad, dyd
        pick = lambda dw: torch.stack([dw[a, c, q, o] for a, c, q, o in idx])
        yb, dxb, dwb = step(xb, wb, dyb, torch.float32)
        yf, dxf, dwf = step(x32, w32, dy32, torch.float32)
        errs = {
            "bf16_rel_l2_vs_fp64": {"fwd": rel(yb[:n], ry), "bwd_input": rel(dxb[:n], rdx),
                                    "bwd_kernel": rel(pick(dwb), rdw)},
            "fp32_rel_l2_vs_fp64": {"fwd": rel(yf[:n], ry), "bwd_input": rel(dxf[:n], rdx),
                                    "bwd_kernel": rel(pick(dwf), rdw)},
            "reference": "torch FP64 on the FP32 inputs: conv2d / conv2d_input on the first 2 images; "
                         "FP64 dot products over all pixels for 256 sampled dw entries",
            "slice": f"y and dx: first {n} images of the full-size calls; dw: 256 entries of the full-size call",
        }
        del x64, w64, g64, ry, rdx, rdw

    # roofline of the dominant BF16 kernel: the main kernel with the longest
    # average launch, timed by the library's HIP events on its own stream
    # over the timed region
    names = ["fwd", "bwd_input", "bwd_kernel"]
    dom = max(names, key=lambda k: kt_b[k]["avg_ms"])
    kname, kavg = kt_b[dom]["kernel"], kt_b[dom]["avg_ms"]
    traffic = TRAFFIC.get((kname, dom))
    traffic = None if traffic is None else int(traffic)
    roof = {"kernel": f"{kname} ({dom})", "bound": "mfma",
            "achieved": round(FLOP_PASS / (kavg * 1e-3) / 1e12, 2), "peak": PEAK["bf16"] / 1e12,
            "unit": "TFLOP/s", "traffic": traffic, "avg_launch_ms": round(kavg, 4),
            "launches": kt_b[dom]["launches"]}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    if traffic is not None:
        roof["traffic_unit"] = "bytes per launch (HBM, rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0)"
    kernels = {k: {"kernel": v["kernel"], "avg_ms": round(v["avg_ms"], 4), "launches": v["launches"]}
               for k, v in kt_b.items()}

    # the Owl drop-in path: FP32 host (numpy, pageable) buffers through the
    # float32 stubs, each call copying its operands to the device, running
    # the same kernels and copying the result back.  Reported beside the
    # device-resident value, never as it.
    xh, wh, dyh = x32.cpu().numpy(), w32.cpu().numpy(), dy32.cpu().numpy()
    host_step = lambda: (nd.conv2d(xh, wh, [1, 1], SAME), nd.conv2d_backward_input(xh, wh, [1, 1], dyh),
                         nd.conv2d_backward_kernel(xh, wh, [1, 1], dyh))
    host_step()
    th = []
    for _ in range(5):
        t0 = time.perf_counter()
        host_step()
        th.append(time.perf_counter() - t0)
    th_med = statistics.median(th)
    pcie_bytes = 4 * (3 * B * H * W * CI + 3 * B * H * W * CO + 3 * K * K * CI * CO)  # per step, both directions
    dropin = {"step_ms": round(th_med * 1e3, 2), "tflops": round(3 * FLOP_PASS / th_med / 1e12, 2),
              "host_device_bytes_per_step": pcie_bytes,
              "what": "Owl float32 drop-ins on pageable host buffers (x, w, dy in; y, dx, dw out): H2D copies + "
                      "FP32 kernels + D2H copies per call, median of 5 steps"}
    del xh, wh, dyh

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()

    out = {
        "metric": "conv fwd+bwd TFLOP/s & % MFMA peak, BF16 vs FP32 plain, at rel-err <=1e-2",
        "value": round(tf_b, 2),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "spinup_s": SPINUP_S,
        "ms_per_step": round(wall_b / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded N(0,1) activations, He-normal kernel)",
        "config": {"workload": "Owl conv2d VGG layer fwd+bwd_input+bwd_kernel, x[64,56,56,64] "
                               "w[3,3,64,128] stride 1 SAME, BF16 in/out (dw FP32), FP32 accumulate",
                   "gflop_per_step": round(3 * FLOP_PASS / 1e9, 3)},
        "bf16_frac_of_peak": round(tf_b * 1e12 / PEAK["bf16"], 4),
        "fp32": {"value": round(tf_f, 2), "ms_per_step": round(wall_f / args.steps * 1e3, 4),
                 "frac_of_peak": round(tf_f * 1e12 / PEAK["f32"], 4)},
        "bf16_over_fp32": round(tf_b / tf_f, 2),
        "per_pass_ms": {"bf16": dict(zip(names, [round(v, 4) for v in per_b])),
                        "fp32": dict(zip(names, [round(v, 4) for v in per_f]))},
        "errors": errs,
        "pool_max2x2s2": pool,
        "conv_bias_relu_bf16": fused,
        "conv_bias_relu_maxpool_bf16": fused_pool,
        "cgraph_vgg_block_bf16": cgraph_block,
        "dropin_fp32_host_buffers": dropin,
        "fp8_e4m3_fwd": fp8,
        "kernels_bf16": kernels,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(budget_s=10.0, budget_1t_s=10.0):
    """Owl's FP32 CPU algorithm as the oracle restates it: for this layer
    (im2col 200704 x 576 >= 512 Ki elements) the cache-blocked GEBP loops
    with Owl's block sizes from the host's CPUID caches, serial packing and
    scatter, and OpenBLAS SGEMM per block.  Whole layer (64 images) on all
    host threads, full steps repeated for about budget_s seconds, median;
    then one BLAS thread on an 8-image sample (still the GEBP branch) for
    about budget_1t_s seconds, median, in TFLOP/s of that sample."""
    import ctypes
    import oracle
    rng = np.random.default_rng(0)
    x = rng.standard_normal([B, H, W, CI]).astype(np.float32)
    w = (rng.standard_normal([K, K, CI, CO]) * (2.0 / (K * K * CI)) ** 0.5).astype(np.float32)
    dy = rng.standard_normal([B, H, W, CO]).astype(np.float32)
    lib, _, path = oracle.openblas_sgemm()
    get_t = getattr(lib, "scipy_openblas_get_num_threads", None) or getattr(lib, "openblas_get_num_threads", None)
    set_t = getattr(lib, "scipy_openblas_set_num_threads", None) or getattr(lib, "openblas_set_num_threads", None)
    cfg = getattr(lib, "scipy_openblas_get_config", None) or getattr(lib, "openblas_get_config", None)
    blas = None
    if cfg is not None:
        cfg.restype = ctypes.c_char_p
        blas = cfg().decode(errors="replace")
    threads = int(get_t()) if get_t is not None else os.cpu_count()

    def bench(nimg, budget):
        g = oracle.conv_geom([nimg, H, W, CI], [K, K, CI, CO], [H, W], [1, 1], [1, 1], [1, 1])
        xs, dys = x[:nimg], dy[:nimg]
        br = []

        def run():
            oracle.conv_fwd_owl_f32(g, xs, w, None, br)
            oracle.conv_bwd_input_owl_f32(g, w, dys, None, br)
            oracle.conv_bwd_kernel_owl_f32(g, xs, dys, None, br)
        run()
        ts = []
        t_start = time.perf_counter()
        while time.perf_counter() - t_start < budget or len(ts) < 3:
            t0 = time.perf_counter()
            run()
            ts.append(time.perf_counter() - t0)
        t = statistics.median(ts)
        flop = 3 * FLOP_PASS * nimg / B
        return flop / t / 1e12, t, len(ts), sum(ts), sorted(set(br))

    tf, t, n, tot, br = bench(B, budget_s)
    one = None
    if set_t is not None:
        set_t(1)
        try:
            tf1, t1, n1, tot1, br1 = bench(8, budget_1t_s)
        finally:
            set_t(threads)
        one = {"value": round(tf1, 5), "unit": "TFLOP/s", "cores": 1, "seconds_per_step": round(t1, 3),
               "sample": f"8 images of the layer ({'/'.join(br1)} branch), fwd+bwd_input+bwd_kernel, "
                         f"{n1} steps in {tot1:.1f} s, median step"}
    return {"value": round(tf, 4), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"the whole layer (64 images), fwd+bwd_input+bwd_kernel, Owl's FP32 {'/'.join(br)} branch "
                      f"(owl_ndarray_conv_impl.h:340-452, 553-663, 774-887; block sizes from CPUID caches "
                      f"{oracle.host_cache_sizes()}) with OpenBLAS sgemm ({os.path.basename(path)}), "
                      f"{n} steps in {tot:.1f} s, median step",
            "seconds_per_step": round(t, 3), "one_thread": one, "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "blas": blas}


if __name__ == "__main__":
    main()
