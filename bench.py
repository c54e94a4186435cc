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
    ("conv_halo_kernel", "fwd"): 13242.9 * 2048 + 50186.2 * 1024,
    ("conv_halo_kernel", "bwd_input"): 25808.6 * 2048 + 25088.1 * 1024,
    ("conv_wgrad_kernel", "bwd_kernel"): 41098.6 * 2048 + 24521.4 * 1024,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 error checks")
    ap.add_argument("--no-configs", action="store_true", help="skip the other workloads (ResNet stack, ragged "
                    "batch, conv3d, small-layer latency)")
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

    configs = None
    if not args.no_configs:
        configs = {"config1_latency": bench_config1(args.steps),
                   "config3_resnet_stack": bench_config3(args.steps, check=not args.no_check),
                   "config4_ragged": bench_config4(args.steps),
                   "config5_conv3d": bench_config5(args.steps)}

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
        "configs": configs,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def _ev_time(f, reps, warm=3):
    """Average milliseconds of f() over reps back-to-back calls between two
    events on the current stream, after warm untimed calls."""
    import torch
    for _ in range(warm):
        f()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        f()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _kernel_avgs(f, reps):
    """The library's per-kernel timing (HIP events stamped with each main
    kernel's own start and end) over reps calls of f: {class: (avg ms, name)}."""
    import torch
    from owl_amd import _lib
    _lib.kernel_timing(True)
    for _ in range(reps):
        f()
    torch.cuda.synchronize()
    out = {}
    for cls, tag in ((_lib.KT_FWD, "fwd"), (_lib.KT_BWD_INPUT, "bwd_input"), (_lib.KT_BWD_KERNEL, "bwd_kernel"),
                     (_lib.KT_AUX, "aux")):
        ms, n, name = _lib.kernel_time(cls)
        if n:
            out[tag] = (ms / n, name)
    _lib.kernel_timing(False)
    return out


def _rel(a, b):
    import torch
    a, b = a.double(), b.double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def _maxabs(a, b):
    return float((a.double() - b.double()).abs().max())


def _dw_samples(x, dy, ksp, stride, pads, idx):
    """FP64 weight-gradient entries idx [(a, c, q, o)] of a 2-D convolution
    (rate 1), each a dot product over every output pixel, on the device."""
    import torch.nn.functional as F
    kh, kw = ksp
    sh, sw = stride
    Ho, Wo = dy.shape[1:3]
    H, W = x.shape[1:3]
    ph, pw = pads
    bot = max(0, sh * (Ho - 1) + kh - H - ph)
    rgt = max(0, sw * (Wo - 1) + kw - W - pw)
    xp = F.pad(x.double(), (0, 0, pw, rgt, ph, bot))
    g = dy.double()
    import torch
    return torch.stack([(xp[:, a:a + sh * (Ho - 1) + 1:sh, c:c + sw * (Wo - 1) + 1:sw, q] * g[..., o]).sum()
                        for a, c, q, o in idx])


def bench_config1(steps):
    """[8,32,32,3] (*) [3,3,3,16] SAME, uniform[-1,1], FP32: latency of the
    device call and of the Owl drop-in on host buffers."""
    import torch
    from owl_amd import ndarray as nd
    from owl_amd.shapes import SAME
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    x = torch.rand(8, 32, 32, 3, device=dev, generator=g) * 2 - 1
    w = torch.rand(3, 3, 3, 16, device=dev, generator=g) * 2 - 1
    t = _ev_time(lambda: nd.conv2d(x, w, [1, 1], SAME), max(steps, 50))
    xh, wh = x.cpu().numpy(), w.cpu().numpy()
    nd.conv2d(xh, wh, [1, 1], SAME)
    hs = []
    for _ in range(20):
        t0 = time.perf_counter()
        nd.conv2d(xh, wh, [1, 1], SAME)
        hs.append(time.perf_counter() - t0)
    flop = 2.0 * 8 * 32 * 32 * 16 * 27
    return {"fwd_device_us": round(t * 1e3, 2), "tflops": round(flop / (t * 1e-3) / 1e12, 3),
            "dropin_host_us": round(statistics.median(hs) * 1e6, 1),
            "what": "FP32 conv2d [8,32,32,3] (*) [3,3,3,16] SAME: device call (events, back to back) and the "
                    "float32 drop-in on host buffers (copies in and out, median of 20 calls)"}


def bench_config3(steps, check=True):
    """ResNet-style stack on [128,28,28,256]: four 3x3 256->256 stride-1 SAME
    convolutions and one 1x1 256->512 stride-2, fwd + bwd_input + bwd_kernel
    of every layer (1440.3 GFLOP per step) in FP32, BF16 and FP16 (FP32
    accumulate), and the FP8 e4m3 forward with per-tensor amax scales
    (480.1 GFLOP).  Errors per layer and pass on the layer's own stored
    inputs against FP64 (torch on the device): y and dx on 2 images, dw on
    64 sampled entries over every pixel; the forward chain against FP64 from
    the FP32 input."""
    import ctypes
    import torch
    import torch.nn.functional as F
    from owl_amd import _lib
    from owl_amd import ndarray as nd
    from owl_amd.shapes import SAME
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(33)
    Bc, Hc, C3 = 128, 28, 256
    x32 = torch.randn(Bc, Hc, Hc, C3, device=dev, generator=g)
    ws32 = [torch.randn(3, 3, C3, C3, device=dev, generator=g) * (2.0 / (9 * C3)) ** 0.5 for _ in range(4)]
    ws32.append(torch.randn(1, 1, C3, 512, device=dev, generator=g) * (2.0 / C3) ** 0.5)
    strides = [[1, 1]] * 4 + [[2, 2]]
    dy_last32 = torch.randn(Bc, 14, 14, 512, device=dev, generator=g)
    flops = [2.0 * Bc * Hc * Hc * C3 * 9 * C3] * 4 + [2.0 * Bc * 14 * 14 * 512 * C3]
    step_flop = 3 * sum(flops)

    def stack(x, ws, dy_last, dt_w):
        acts = [x]
        for w, st in zip(ws, strides):
            acts.append(nd.conv2d(acts[-1], w, st, SAME))
        dy = dy_last
        dws, dys = [None] * 5, [None] * 5
        for l in range(4, -1, -1):
            dys[l] = dy
            dws[l] = nd.conv2d_backward_kernel(acts[l], ws[l], strides[l], dy, out_dtype=dt_w)
            dy = nd.conv2d_backward_input(acts[l], ws[l], strides[l], dy)
        return acts, dys, dws, dy  # dy: the input's gradient

    out = {"workload": "[128,28,28,256]: 4 x conv 3x3 256->256 s1 SAME + conv 1x1 256->512 s2 SAME, "
                       "fwd+bwd_input+bwd_kernel per layer", "gflop_per_step": round(step_flop / 1e9, 1)}
    gi = torch.Generator()
    gi.manual_seed(5)
    for name, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16), ("fp16", torch.float16)):
        x, ws, dyl = x32.to(dt), [w.to(dt) for w in ws32], dy_last32.to(dt)
        run = lambda: stack(x, ws, dyl, torch.float32)
        n = max(3, steps // (4 if dt == torch.float32 else 1))
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            run()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / n * 1e3
        peak = PEAK["f32"] if dt == torch.float32 else PEAK["bf16"]
        rec = {"ms_per_step": round(ms, 3), "tflops": round(step_flop / (ms * 1e-3) / 1e12, 1),
               "frac_of_peak": round(step_flop / (ms * 1e-3) / peak, 4)}
        # the dominant kernel of a 3x3 layer's three passes
        a0 = x
        kt = _kernel_avgs(lambda: (nd.conv2d(a0, ws[0], [1, 1], SAME),
                                   nd.conv2d_backward_input(a0, ws[0], [1, 1], a0),
                                   nd.conv2d_backward_kernel(a0, ws[0], [1, 1], a0, out_dtype=torch.float32)), 5)
        dom = max(("fwd", "bwd_input", "bwd_kernel"), key=lambda k: kt[k][0])
        rec["roofline_3x3"] = {"kernel": f"{kt[dom][1]} ({dom})", "avg_launch_ms": round(kt[dom][0], 4),
                               "achieved": round(flops[0] / (kt[dom][0] * 1e-3) / 1e12, 1),
                               "frac": round(flops[0] / (kt[dom][0] * 1e-3) / peak, 4),
                               "kernels_ms": {k: round(v[0], 4) for k, v in kt.items()}}
        if check:
            acts, dys, dws, dx0 = run()
            torch.cuda.synchronize()
            errs = []
            ref = x32[:2].double()
            for l in range(5):
                st, w64 = strides[l], ws[l].double().permute(3, 2, 0, 1).contiguous()
                pad = 1 if l < 4 else 0
                a_in = acts[l][:2].double().permute(0, 3, 1, 2)
                # SAME pads 1 for the 3x3 stride-1 layers, nothing for the 1x1 stride-2 one
                ry = F.conv2d(a_in, w64, stride=st, padding=pad).permute(0, 2, 3, 1)
                g64 = dys[l][:2].double().permute(0, 3, 1, 2)
                rdx = torch.nn.grad.conv2d_input(a_in.shape, w64, g64, stride=st, padding=pad).permute(0, 2, 3, 1)
                ks = list(ws[l].shape)
                flat = torch.randperm(ks[0] * ks[1] * ks[2] * ks[3], generator=gi)[:64]
                idx = torch.stack([flat // (ks[1] * ks[2] * ks[3]), flat // (ks[2] * ks[3]) % ks[1],
                                   flat // ks[3] % ks[2], flat % ks[3]], 1).tolist()
                rdw = _dw_samples(acts[l], dys[l], ks[:2], st, [pad, pad], idx)
                gdw = torch.stack([dws[l][a, c, q, o] for a, c, q, o in idx])
                ref = F.conv2d(ref.permute(0, 3, 1, 2), ws32[l].double().permute(3, 2, 0, 1).contiguous(),
                               stride=st, padding=pad).permute(0, 2, 3, 1)
                dx = (dys[l - 1] if l > 0 else dx0)[:2]
                errs.append({"layer": l, "fwd_rel_l2": _rel(acts[l + 1][:2], ry), "fwd_max_abs": _maxabs(acts[l + 1][:2], ry),
                             "bwd_input_rel_l2": _rel(dx, rdx), "bwd_input_max_abs": _maxabs(dx, rdx),
                             "bwd_kernel_rel_l2_sampled": _rel(gdw, rdw), "bwd_kernel_max_abs_sampled": _maxabs(gdw, rdw),
                             "fwd_chain_rel_l2_vs_fp32_input": _rel(acts[l + 1][:2], ref)})
            rec["errors"] = errs
        out[name] = rec
        del x, ws, dyl
    # FP8 e4m3 forward of the stack: each layer quantises its BF16 input and
    # its weights (per-tensor amax scales), output BF16
    st_ = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = _lib.lib()

    def quant(t):
        q = torch.empty(t.numel(), dtype=torch.uint8, device=dev)
        sc = torch.empty(1, dtype=torch.float32, device=dev)
        _lib.check(L.owl_amd_fp8_quantize(_lib.BF16, ctypes.c_void_p(t.data_ptr()), t.numel(),
                                          ctypes.c_void_p(q.data_ptr()), ctypes.c_void_p(sc.data_ptr()), st_))
        return q, sc
    wq = [quant(w.bfloat16().contiguous()) for w in ws32]
    geos = []
    shp = [Bc, Hc, Hc, C3]
    for w, stv in zip(ws32, strides):
        geo = nd.conv_problem(shp, list(w.shape), stv, None, SAME)
        geos.append(geo)
        shp = [Bc, *geo.sp_out, w.shape[3]]
    ys8 = [torch.empty([Bc, *g_.sp_out, w.shape[3]], dtype=torch.bfloat16, device=dev) for g_, w in zip(geos, ws32)]
    descs = [g_.desc() for g_ in geos]

    def fp8_stack():
        a = x32.bfloat16()
        for l in range(5):
            q, sc = quant(a)
            _lib.check(L.owl_amd_conv_fwd_fp8(ctypes.byref(descs[l]), _lib.BF16, ctypes.c_void_p(q.data_ptr()),
                                              ctypes.c_void_p(wq[l][0].data_ptr()), ctypes.c_void_p(sc.data_ptr()),
                                              ctypes.c_void_p(wq[l][1].data_ptr()),
                                              ctypes.c_void_p(ys8[l].data_ptr()), st_))
            a = ys8[l]
    fp8_stack()
    torch.cuda.synchronize()
    n = max(3, steps)
    t0 = time.perf_counter()
    for _ in range(n):
        fp8_stack()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / n * 1e3
    fwd_flop = sum(flops)
    rec = {"ms_per_fwd": round(ms, 3), "tflops": round(fwd_flop / (ms * 1e-3) / 1e12, 1),
           "frac_of_fp8_peak": round(fwd_flop / (ms * 1e-3) / 5.0e15, 4),
           "what": "forward of the stack, per-layer quantisation of the BF16 input included"}
    if check:
        errs = []
        a_prev = x32.bfloat16()
        for l in range(5):
            pad = 1 if l < 4 else 0
            ry = F.conv2d(a_prev[:2].double().permute(0, 3, 1, 2), ws32[l].double().permute(3, 2, 0, 1).contiguous(),
                          stride=strides[l], padding=pad).permute(0, 2, 3, 1)
            errs.append({"layer": l, "fwd_rel_l2": _rel(ys8[l][:2], ry), "fwd_max_abs": _maxabs(ys8[l][:2], ry)})
            a_prev = ys8[l]
        rec["errors_vs_fp64_on_bf16_input"] = errs
    out["fp8_e4m3_fwd"] = rec
    return out


def bench_config4(steps):
    """64 seeded ragged conv2d problems (BF16, FP32 accumulate): one
    persistent grouped launch (autotuned per shape class: tile edge and
    split-K, or a launch of its own), its hipGraph replay, and the same 64
    problems as separate launches."""
    import torch
    from owl_amd import ndarray as nd
    from owl_amd.group import ConvGroup, ragged_problems
    dev = torch.device("cuda:0")
    problems = ragged_problems(seed=4, n=64)
    g = torch.Generator(device=dev)
    g.manual_seed(44)
    xs = [(torch.rand(p[0], device=dev, generator=g) * 2 - 1).bfloat16() for p in problems]
    ws = [(torch.rand(p[1], device=dev, generator=g) * 2 - 1).bfloat16() for p in problems]
    grp = ConvGroup(problems, torch.bfloat16, torch.bfloat16)
    ys = grp.empty_outputs(dev)
    grp.autotune(xs, ws, ys)
    flop = 0.0
    for geo in grp.geoms:
        flop += 2.0 * geo.B * int(np.prod(geo.sp_out)) * geo.Co * int(np.prod(geo.ksp)) * geo.Ci
    reps = max(steps, 10)
    t_group = _ev_time(lambda: grp.run(xs, ws, ys), reps)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    grp.capture(xs, ws, ys, s)
    torch.cuda.synchronize()
    t_graph = _ev_time(lambda: grp.replay(torch.cuda.current_stream()), reps)
    sep = lambda: [nd.dilated_conv2d(x, w, p[2], p[3], p[4]) for x, w, p in zip(xs, ws, problems)]
    t_sep = _ev_time(sep, max(3, reps // 2))
    t_sep_wall = None
    sep()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        sep()
    torch.cuda.synchronize()
    t_sep_wall = (time.perf_counter() - t0) / 3 * 1e3
    edges, tiles = grp.tile_edges()
    return {"gflop": round(flop / 1e9, 2), "group_ms": round(t_group, 4), "graph_replay_ms": round(t_graph, 4),
            "separate_launches_ms": round(t_sep, 4), "separate_launches_wall_ms": round(t_sep_wall, 3),
            "group_tflops": round(flop / (t_group * 1e-3) / 1e12, 1),
            "graph_tflops": round(flop / (t_graph * 1e-3) / 1e12, 1),
            "separate_tflops": round(flop / (t_sep * 1e-3) / 1e12, 1),
            "group_frac_of_bf16_peak": round(flop / (t_group * 1e-3) / PEAK["bf16"], 4),
            "tiles": tiles, "tile_edges": sorted(set(e for e in edges if e)),
            "own_launches": sum(1 for e in edges if e == 0),
            "split_k_problems": sum(1 for sp in grp.plan()[1] if sp > 1),
            "what": "64 conv2d problems (seed 4): B 1-32, H=W 7-112, Ci/Co 16-512, k 1/3/5, s 1/2, rate 1/2, "
                    "SAME/VALID; one persistent launch (tile edge and split-K per shape class, plus the classes "
                    "the tuner gives a launch of their own) vs the same calls one by one"}


def bench_config5(steps):
    """conv3d [8,32,32,32,16] (*) [3,3,3,16,32] (*) [3,3,3,32,64] SAME, then
    avg_pool3d 2x2x2: fwd + bwd_input + bwd_kernel of both convolutions
    (108.7 GFLOP) in FP32 and BF16, and the average pool's GB/s."""
    import torch
    from owl_amd import ndarray as nd
    from owl_amd.shapes import SAME, VALID
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(55)
    x32 = torch.randn(8, 32, 32, 32, 16, device=dev, generator=g)
    w1 = torch.randn(3, 3, 3, 16, 32, device=dev, generator=g) * (2.0 / 432) ** 0.5
    w2 = torch.randn(3, 3, 3, 32, 64, device=dev, generator=g) * (2.0 / 864) ** 0.5
    flop = 3 * 2.0 * 8 * 32 ** 3 * 27 * (16 * 32 + 32 * 64)
    s1 = [1, 1, 1]
    out = {"gflop_per_step": round(flop / 1e9, 2)}
    for name, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16)):
        x, a, b = x32.to(dt), w1.to(dt), w2.to(dt)
        y1 = nd.conv3d(x, a, s1, SAME)
        y2 = nd.conv3d(y1, b, s1, SAME)
        dy2 = torch.randn(y2.shape, device=dev, generator=g).to(dt)

        def step():
            h1 = nd.conv3d(x, a, s1, SAME)
            h2 = nd.conv3d(h1, b, s1, SAME)
            nd.conv3d_backward_kernel(h1, b, s1, dy2, out_dtype=torch.float32)
            d1 = nd.conv3d_backward_input(h1, b, s1, dy2)
            nd.conv3d_backward_kernel(x, a, s1, d1, out_dtype=torch.float32)
            nd.conv3d_backward_input(x, a, s1, d1)
            return h2
        ms = _ev_time(step, max(3, steps // 2))
        peak = PEAK["f32"] if dt == torch.float32 else PEAK["bf16"]
        kt = _kernel_avgs(lambda: (nd.conv3d(y1, b, s1, SAME), nd.conv3d_backward_input(y1, b, s1, dy2),
                                   nd.conv3d_backward_kernel(y1, b, s1, dy2, out_dtype=torch.float32)), 3)
        f2 = 2.0 * 8 * 32 ** 3 * 27 * 32 * 64
        dom = max(("fwd", "bwd_input", "bwd_kernel"), key=lambda k: kt[k][0])
        out[name] = {"ms_per_step": round(ms, 3), "tflops": round(flop / (ms * 1e-3) / 1e12, 1),
                     "frac_of_peak": round(flop / (ms * 1e-3) / peak, 4),
                     "roofline_layer2": {"kernel": f"{kt[dom][1]} ({dom})", "avg_launch_ms": round(kt[dom][0], 4),
                                         "frac": round(f2 / (kt[dom][0] * 1e-3) / peak, 4)}}
    y2 = nd.conv3d(nd.conv3d(x32, w1, s1, SAME), w2, s1, SAME)
    pool_f = _ev_time(lambda: nd.avg_pool3d(y2, [2, 2, 2], [2, 2, 2], VALID), 20)
    p = nd.avg_pool3d(y2, [2, 2, 2], [2, 2, 2], VALID)
    gp = torch.randn(p.shape, device=dev, generator=g)
    pool_b = _ev_time(lambda: nd.avg_pool3d_backward(VALID, y2, [2, 2, 2], [2, 2, 2], gp), 20)
    nin, nout = y2.numel() * 4, p.numel() * 4
    out["avg_pool3d_fp32"] = {"fwd_GBps": round((nin + nout) / (pool_f * 1e-3) / 1e9, 1),
                              "bwd_GBps": round((nout + nin) / (pool_b * 1e-3) / 1e9, 1),
                              "fwd_frac_hbm": round((nin + nout) / (pool_f * 1e-3) / HBM_PEAK, 3),
                              "bwd_frac_hbm": round((nout + nin) / (pool_b * 1e-3) / HBM_PEAK, 3)}
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_leg(nimg, budget):
    """Owl's FP32 CPU algorithm (the oracle's restatement) on nimg images of
    the layer, full steps repeated for about budget seconds: (TFLOP/s, median
    seconds per step, steps, total seconds, branches taken)."""
    import oracle
    rng = np.random.default_rng(0)
    x = rng.standard_normal([nimg, H, W, CI]).astype(np.float32)
    w = (rng.standard_normal([K, K, CI, CO]) * (2.0 / (K * K * CI)) ** 0.5).astype(np.float32)
    dy = rng.standard_normal([nimg, H, W, CO]).astype(np.float32)
    g = oracle.conv_geom([nimg, H, W, CI], [K, K, CI, CO], [H, W], [1, 1], [1, 1], [1, 1])
    br = []

    def run():
        oracle.conv_fwd_owl_f32(g, x, w, None, br)
        oracle.conv_bwd_input_owl_f32(g, w, dy, None, br)
        oracle.conv_bwd_kernel_owl_f32(g, x, dy, None, br)
    run()
    ts = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget or len(ts) < 3:
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    t = statistics.median(ts)
    return 3 * FLOP_PASS * nimg / B / t / 1e12, t, len(ts), sum(ts), sorted(set(br))


def cpu_baseline(budget_s=10.0, budget_1t_s=10.0):
    """Owl's FP32 CPU algorithm as the oracle restates it: for this layer
    (im2col 200704 x 576 >= 512 Ki elements) the cache-blocked GEBP loops
    with Owl's block sizes from the host's CPUID caches, serial packing and
    scatter, and OpenBLAS SGEMM per block.  Whole layer (64 images) on all
    host threads, full steps repeated for about budget_s seconds, median;
    then, in a child process started with OPENBLAS_NUM_THREADS=1, one BLAS
    thread on an 8-image sample (still the GEBP branch) for about
    budget_1t_s seconds, median, in TFLOP/s of that sample."""
    import ctypes
    import subprocess
    import oracle
    lib, _, path = oracle.openblas_sgemm()
    get_t = getattr(lib, "scipy_openblas_get_num_threads", None) or getattr(lib, "openblas_get_num_threads", None)
    cfg = getattr(lib, "scipy_openblas_get_config", None) or getattr(lib, "openblas_get_config", None)
    blas = None
    if cfg is not None:
        cfg.restype = ctypes.c_char_p
        blas = cfg().decode(errors="replace")
    threads = int(get_t()) if get_t is not None else os.cpu_count()
    tf, t, n, tot, br = _cpu_leg(B, budget_s)
    one = None
    env = dict(os.environ, OPENBLAS_NUM_THREADS="1", OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", "8", str(budget_1t_s)],
                       capture_output=True, text=True, env=env, timeout=budget_1t_s * 10 + 60)
    if r.returncode == 0:
        leg = json.loads(r.stdout.strip().splitlines()[-1])
        one = {"value": round(leg["tflops"], 5), "unit": "TFLOP/s", "cores": leg["threads"],
               "seconds_per_step": round(leg["seconds"], 3),
               "sample": f"8 images of the layer ({'/'.join(leg['branches'])} branch), fwd+bwd_input+bwd_kernel, "
                         f"{leg['steps']} steps in {leg['total']:.1f} s, median step (a child process with "
                         f"OPENBLAS_NUM_THREADS=1)"}
    return {"value": round(tf, 4), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"the whole layer (64 images), fwd+bwd_input+bwd_kernel, Owl's FP32 {'/'.join(br)} branch "
                      f"(owl_ndarray_conv_impl.h:340-452, 553-663, 774-887; block sizes from CPUID caches "
                      f"{oracle.host_cache_sizes()}) with OpenBLAS sgemm ({os.path.basename(path)}), "
                      f"{n} steps in {tot:.1f} s, median step",
            "seconds_per_step": round(t, 3), "one_thread": one, "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "blas": blas}


def cpu_leg_main(argv):
    """Child-process entry of the one-thread CPU leg: prints one JSON line."""
    import oracle
    nimg, budget = int(argv[0]), float(argv[1])
    lib, _, _ = oracle.openblas_sgemm()
    get_t = getattr(lib, "scipy_openblas_get_num_threads", None) or getattr(lib, "openblas_get_num_threads", None)
    tf, t, n, tot, br = _cpu_leg(nimg, budget)
    print(json.dumps({"tflops": tf, "seconds": t, "steps": n, "total": tot, "branches": br,
                      "threads": int(get_t()) if get_t is not None else None}))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-leg":
        cpu_leg_main(sys.argv[2:])
    else:
        main()
