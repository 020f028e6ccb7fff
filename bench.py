#!/usr/bin/env python3
"""Benchmark: fp16 batched 1D FFT, N = 2^16, batch 256 (the headline workload).

One step = one tcfftExec over the whole batch (256 signals of 65536 complex
fp16 values, in place, digit-reversed input, natural output) on one MI355X.
Inputs are synthetic uniform[-1,1) fp16 and already resident in HBM; the timed
steps rotate over ROTATE buffers (> 512 MiB in total) so no step starts with
its data in the 256 MB MALL.

Prints one JSON line:
  value      = 5*N*log2(N)*batch / step time, GFLOP/s
  roofline   = dominant pass kernel: algorithmic bytes per launch
               (8 B per complex element: one fp16 read + one fp16 write)
               / its mean launch time (HIP events on the launch stream),
               against 8 TB/s; traffic from rocprofv3 counters when a
               profile summary is committed, else null
  cpu_baseline = the FP64 checker (oracle/, a scalar C restatement of the
               reference algorithm) on a bounded sample, 1 host core
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 1 << 16
BATCH = 256
ROTATE = 10          # 10 x 64 MiB = 640 MiB
PEAK_HBM_GBS = 8000.0
METRIC = "fp16 FFT GFLOP/s (5N·log2N) + % HBM roofline at N=2^16; rel-L2 vs FP64"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=N)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="plain", choices=["plain", "split", "uneven"],
                    help="plain: the headline fp16 exec; split: FP32 in/out split-precision exec "
                         "(reported with its time ratio to the plain exec)")
    ap.add_argument("--split-chunk", type=int, default=None, help="signals per split pass pair")
    ap.add_argument("--calib", action="store_true",
                    help="also run 5 device-to-device copies of one 64 MiB buffer (counter calibration kernel)")
    return ap.parse_args()


def cpu_baseline(n: int):
    """FP64 checker on a bounded sample of the same workload (1 core)."""
    import oracle
    oracle.lib()
    rng = np.random.default_rng(0)
    sig = 4
    x = (rng.uniform(-1, 1, (sig, n)) + 1j * rng.uniform(-1, 1, (sig, n))).astype(np.complex64).astype(np.complex128)
    reps = 0
    t0 = time.perf_counter()
    while True:
        oracle.fft(x)
        reps += 1
        el = time.perf_counter() - t0
        if el > 10.0 or reps >= 50:
            break
    flops = 5.0 * n * math.log2(n) * sig * reps
    return {"value": round(flops / el / 1e9, 3), "unit": "GFLOP/s", "cores": 1, "kind": "port",
            "sample": f"{reps} x {sig} signals of N={n} (FP64 DIT restatement, {el:.1f} s)"}


def load_traffic(n: int, batch: int):
    """Counter-derived HBM bytes per launch of the dominant kernel, if profiled."""
    path = os.path.join(ROOT, "profiles", f"traffic_n{n}_b{batch}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def graph_time(torch, launch, reps):
    """Mean ms per launch(i, stream) over reps, captured once into a hipGraph
    and replayed between synchronizations."""
    g = torch.cuda.CUDAGraph()
    cs = torch.cuda.Stream()
    cs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=cs):
        for i in range(reps):
            launch(i, cs.cuda_stream)
    g.replay()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(torch.cuda.current_stream())
    g.replay()
    b.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def bench_split(args):
    """Split precision, N = 2^16, FP32 complex in/out, N(0,1)*10^U(-3,3) per
    signal; reports the time ratio to the plain fp16 exec of the same shape."""
    import torch
    import tcfft_amd
    n, batch = 1 << 16, args.batch
    rot = 4  # 4 x (128 MiB in + 128 MiB out) > MALL
    rng = np.random.default_rng(0)
    scale = 10.0 ** rng.uniform(-3, 3, (batch, 1))
    x = ((rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))) * scale).astype(np.complex64)
    host = torch.from_numpy(x.view(np.float32))
    ins = [host.cuda() for _ in range(rot)]
    outs = [torch.empty_like(ins[0]) for _ in range(rot)]
    plan = tcfft_amd.Plan1D(n, batch)
    if args.split_chunk is not None:
        plan.set_option(tcfft_amd.OPT_SPLIT_CHUNK, args.split_chunk)
    plan.exec_split(ins[0].data_ptr(), outs[0].data_ptr())
    torch.cuda.synchronize()
    k = 8
    got = outs[0][:k].cpu().numpy().view(np.complex64).astype(np.complex128)
    ref = np.fft.fft(x[:k].astype(np.complex128), axis=-1)
    errs = [float(np.linalg.norm(got[i] - ref[i]) / np.linalg.norm(ref[i])) for i in range(k)]
    for i in range(args.warmup):
        plan.exec_split(ins[i % rot].data_ptr(), outs[i % rot].data_ptr())
    ms = graph_time(torch, lambda i, st: plan.exec_split(ins[i % rot].data_ptr(), outs[i % rot].data_ptr(),
                                                         tcfft_amd.FORWARD, st), args.steps)
    # plain fp16 exec of the same shape, same run
    pbufs = [torch.empty((batch, n, 2), dtype=torch.float16, device="cuda") for _ in range(10)]
    for b in pbufs:
        b.uniform_(-1, 1)
    ms_plain = graph_time(torch, lambda i, st: plan.exec(pbufs[i % 10].data_ptr(), tcfft_amd.FORWARD, st), args.steps)
    flops = 5.0 * n * math.log2(n) * batch
    out = {"metric": "split-precision FFT GFLOP/s (5N·log2N) at N=2^16; rel-L2 vs FP64 per signal",
           "value": round(flops / (ms * 1e-3) / 1e9, 1), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f16 hi+lo MFMA, FP32 accumulate, FP32 I/O",
           "data": "synthetic N(0,1)*10^U(-3,3) per signal, complex64",
           "config": {"workload": f"1D C2C split-precision FFT N={n} batch={batch}, FP32 in/out, natural order",
                      "n": n, "batch": batch, "split_chunk": args.split_chunk},
           "rel_l2_max_per_signal": max(errs), "ms_plain_fp16": round(ms_plain, 5),
           "time_ratio_vs_plain": round(ms / ms_plain, 3),
           "alg_bytes": 16 * n * batch, "hbm_frac_alg": round(16 * n * batch / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    plan.destroy()
    print(json.dumps(out), flush=True)


def bench_uneven(args):
    """4096 fp16 signals, N = 2^U{8..16} (seeded), drop-in order, one
    persistent autotuned launch vs one launch per size class (regular plans on
    a size-sorted copy of the same signals)."""
    import torch
    import tcfft_amd
    rng = np.random.default_rng(0)
    n_sig = 4096
    sizes = (2 ** rng.integers(8, 17, n_sig)).astype(np.int32)
    total = int(sizes.sum())
    flat = rng.uniform(-1, 1, (total, 2)).astype(np.float16)  # contents are synthetic per signal
    mp = tcfft_amd.MultiPlan(sizes)
    rot = 3  # 3 x 238 MB > MALL
    host = torch.from_numpy(flat)
    bufs = [host.cuda() for _ in range(rot)]
    g, o, tune_ms = mp.autotune(bufs[0].data_ptr())
    for b in bufs:
        b.copy_(host)
    ms_multi = graph_time(torch, lambda i, st: mp.exec(bufs[i % rot].data_ptr(), tcfft_amd.FORWARD, st), args.steps)
    # per-size launches on a size-sorted layout of the same data volume
    classes = sorted(set(int(n) for n in sizes))
    plans, offs, o2 = [], [], 0
    for n in classes:
        cnt = int((sizes == n).sum())
        plans.append(tcfft_amd.Plan1D(n, cnt))
        offs.append(o2)
        o2 += n * cnt

    def per_size(i, st):
        base = bufs[i % rot].data_ptr()
        for p, off in zip(plans, offs):
            p.exec(base + 4 * off, tcfft_amd.FORWARD, st)

    ms_sizes = graph_time(torch, per_size, args.steps)
    # accuracy of the persistent launch on a sample of signals (fresh data)
    sample = [int(i) for i in rng.choice(n_sig, 16, replace=False)]
    check = torch.from_numpy(flat).cuda()
    xs = {}
    for i in sample:
        n, o3 = int(sizes[i]), int(mp.offsets[i])
        rad, _ = tcfft_amd.query_1d(n, 1)
        x = flat[o3:o3 + n].astype(np.float64)
        xs[i] = x[:, 0] + 1j * x[:, 1]
        xrev = flat[o3:o3 + n][tcfft_amd.gen_rev(n, rad)]
        check[o3:o3 + n] = torch.from_numpy(np.ascontiguousarray(xrev)).cuda()
    mp.exec(check.data_ptr())
    torch.cuda.synchronize()
    errs = []
    for i in sample:
        n, o3 = int(sizes[i]), int(mp.offsets[i])
        y = check[o3:o3 + n].float().cpu().numpy().astype(np.float64)
        yc = y[:, 0] + 1j * y[:, 1]
        ref = np.fft.fft(xs[i])
        errs.append(float(np.linalg.norm(yc - ref) / np.linalg.norm(ref)))
    flops = float(sum(5.0 * n * math.log2(n) for n in sizes))
    out = {"metric": "uneven-batch FFT GFLOP/s (sum 5N·log2N), one persistent launch",
           "value": round(flops / (ms_multi * 1e-3) / 1e9, 1), "unit": "GFLOP/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 0, "ms_per_step": round(ms_multi, 5), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f16 (FP32 MFMA accumulate)",
           "data": "synthetic uniform[-1,1) fp16, 4096 signals, N=2^U{8..16} seed 0",
           "config": {"workload": "uneven batch, one persistent hipGraph-captured launch", "signals": n_sig,
                      "elements": total, "tasks": mp.n_tasks, "autotuned_grid": g, "autotuned_order": o},
           "ms_per_size_launches": round(ms_sizes, 5), "speedup_vs_per_size": round(ms_sizes / ms_multi, 3),
           "rel_l2_max_sample": max(errs),
           "hbm_frac_alg": round(8.0 * total / (ms_multi * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    for p in plans:
        p.destroy()
    mp.destroy()
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.workload == "uneven":
        if args.gpus != 1:
            sys.exit(2)
        bench_uneven(args)
        return
    if args.workload == "split":
        if args.gpus != 1:
            sys.exit(2)
        bench_split(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batch axis shards with "
              f"no exchange, see DESIGN.md. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    import tcfft_amd

    n, batch = args.n, args.batch
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    plan = tcfft_amd.Plan1D(n, batch)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    # synthetic fp16 inputs, digit-reversed by the plan's radices (gen_rev)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (batch, n, 2)).astype(np.float16)
    rev = tcfft_amd.gen_rev(n, plan.radices)
    x_rev = np.ascontiguousarray(x[:, rev, :])
    host = torch.from_numpy(x_rev)
    bufs = [torch.empty((batch, n, 2), dtype=torch.float16, device=dev) for _ in range(ROTATE)]
    for b in bufs:
        b.copy_(host)
    torch.cuda.synchronize()

    if args.calib:
        for _ in range(5):
            bufs[1].copy_(bufs[0])
        torch.cuda.synchronize()

    # accuracy of one exec vs FP64 (numpy pocketfft) on a few signals
    check = bufs[0]
    plan.exec(check.data_ptr(), tcfft_amd.FORWARD, sptr)
    torch.cuda.synchronize()
    k = min(batch, 8)
    got = check[:k].float().cpu().numpy().astype(np.float64)
    gotc = got[..., 0] + 1j * got[..., 1]
    xin = x[:k].astype(np.float64)
    ref = np.fft.fft(xin[..., 0] + 1j * xin[..., 1], axis=-1)
    rel_l2 = float(np.linalg.norm(gotc - ref) / np.linalg.norm(ref))
    check.copy_(host)

    def timed(launch, reps, graph):
        """Mean ms per launch() over reps, bracketed by synchronize; with
        graph=True the reps launches are captured once into a hipGraph and
        replayed (no host launch overhead inside the timed region)."""
        if graph:
            g = torch.cuda.CUDAGraph()
            cs = torch.cuda.Stream()
            cs.wait_stream(torch.cuda.current_stream())
            with torch.cuda.graph(g, stream=cs):
                for i in range(reps):
                    launch(i, cs.cuda_stream)
            g.replay()  # untimed replay (replay() runs on the current stream)
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(torch.cuda.current_stream())
            g.replay()
            b.record(torch.cuda.current_stream())
            torch.cuda.synchronize()
            return a.elapsed_time(b) / reps
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for i in range(reps):
            launch(i, sptr)
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    def exec_step(i, st):
        plan.exec(bufs[i % ROTATE].data_ptr(), tcfft_amd.FORWARD, st)

    # warmup (eager), then the timed K steps: eager and graph-replayed
    for i in range(args.warmup):
        exec_step(i, sptr)
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    ms_eager = timed(exec_step, args.steps, graph=False)
    ms = timed(exec_step, args.steps, graph=True)
    t_wall = time.perf_counter() - t_wall0

    # per-pass timing (each pass kernel alone, graph-replayed on its stream)
    npass = plan.num_passes
    per_pass = []
    for p in range(npass):
        def pass_step(i, st, p=p):
            plan.exec_passes(bufs[i % ROTATE].data_ptr(), p, 1, tcfft_amd.FORWARD, st)
        per_pass.append(timed(pass_step, max(20, args.steps), graph=True))
    dom = int(np.argmax(per_pass))
    alg_bytes = 8.0 * n * batch
    achieved = alg_bytes / (per_pass[dom] * 1e-3) / 1e9
    exec_gbs = alg_bytes / (ms * 1e-3) / 1e9

    flops = 5.0 * n * math.log2(n) * batch
    gflops = flops / (ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gflops, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "ms_per_step_eager": round(ms_eager, 5),
        "timing": "hipGraph replay of the K timed steps (eager launch loop in ms_per_step_eager)",
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 (FP32 MFMA accumulate)",
        "data": "synthetic uniform[-1,1) fp16, resident in HBM, rotated over %d buffers" % ROTATE,
        "config": {"workload": f"1D C2C FFT N={n} batch={batch}, fp16 interleaved, in place, "
                               f"digit-reversed input (tcfftExec)", "n": n, "batch": batch,
                   "passes": npass, "radices": plan.radices},
        "rel_l2_vs_fp64": rel_l2,
        "hbm_frac_whole_exec": round(exec_gbs / PEAK_HBM_GBS, 4),
        "pass_ms": [round(v, 5) for v in per_pass],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": load_traffic(n, batch),
                     "kernel": f"pass {dom}"},
        "wall_s_timed": round(t_wall, 3),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n)
    plan.destroy()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
