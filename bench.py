#!/usr/bin/env python3
"""Benchmark: fp16 batched 1D FFT, N = 2^16, batch 256 (the headline workload),
plus the other measured workloads in the same run.

One step = one tcfftExec over the whole batch (256 signals of 65536 complex
fp16 values, in place, digit-reversed input, natural output) on one MI355X.
Inputs are synthetic uniform[-1,1) fp16 and already resident in HBM; the timed
steps rotate over buffers totalling > 640 MiB so no step starts with its data
in the 256 MB MALL.  Steps are captured once into a hipGraph and replayed.
Every in-place workload transforms each buffer at most once per timed region
and gets its input back before the next (Rotation.refresh): repeated in place,
fp16 data overflows to inf/NaN, which runs faster than finite data.

Prints one JSON line:
  value      = 5*N*log2(N)*batch / step time, GFLOP/s
  roofline   = the 65536-point pass kernel: algorithmic bytes per launch
               (8 B per complex element: one fp16 read + one fp16 write)
               / its mean launch time (HIP events on the launch stream),
               against 8 TB/s; traffic = HBM bytes per launch from rocprofv3
               FETCH_SIZE / WRITE_SIZE counters when that summary is committed
  cpu_baseline = the FP64 checker (oracle/, a scalar C restatement of the
               reference algorithm) on a bounded sample, 1 host core
  cpu_fft_f32  = the single-precision host FFT the GPU path replaces, on this
               process's CPU share, bounded sample: FFTW3 API (libfftw3f,
               else MKL's wrappers) in a child process, else scipy.fft
  natural_order = the same signals in natural order through tcfft_exec_natural
               (out of place), its time, HBM fraction and rel-L2
  configs    = the other workloads (N=4096 x 1024; N=2^20 x 16; 2D 1024^2 x 8;
               split precision; uneven batch; FFT-conv layer), each with its
               time, GFLOP/s, HBM fraction and rel-L2 vs FP64, and hipFFT's
               half-precision transform on three of them as the vendor line
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 1 << 16
BATCH = 256
ROTATE_BYTES = 640 << 20   # rotation set > MALL (256 MB)
PEAK_HBM_GBS = 8000.0
METRIC = "fp16 FFT GFLOP/s (5N·log2N) + % HBM roofline at N=2^16; rel-L2 vs FP64"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=N)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-configs", action="store_true", help="headline workload only")
    ap.add_argument("--workload", default="plain", choices=["plain", "split", "uneven", "conv", "r2c"],
                    help="plain: the headline fp16 exec (+ every config); split / uneven / conv: that leg alone")
    ap.add_argument("--split-chunk", type=int, default=None, help="signals per split pass pair")
    ap.add_argument("--calib", action="store_true",
                    help="also run 5 device-to-device copies of one 64 MiB buffer (counter calibration kernel)")
    return ap.parse_args()


# ----------------------------------------------------------------------------
# timing helpers

# GPU cycles of the spin queued before each timed graph replay (about
# 0.1 ms): the replay is submitted while the GPU spins, so the timed region
# (start event -> end event) holds the K steps back to back and not the
# host-side launch latency of the graph, which at K = 20 adds ~1 us per step.
SPIN_CYCLES = 250_000


def graph_time(torch, launch, reps, replays=1, refresh=None):
    """ms per launch(i, stream): reps launches captured once into a hipGraph,
    replayed `replays` times between synchronizations (replay() runs on the
    current stream, so the events are recorded there); median over replays.
    Each replay is queued behind a short spin kernel (SPIN_CYCLES) that ends
    before the start event.  refresh(), if given, runs before every replay,
    outside the timed region (in-place workloads put their input back)."""
    g = torch.cuda.CUDAGraph()
    cs = torch.cuda.Stream()
    cs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=cs):
        for i in range(reps):
            launch(i, cs.cuda_stream)
    if refresh:
        refresh()
    g.replay()
    torch.cuda.synchronize()
    times = []
    for _ in range(replays):
        if refresh:
            refresh()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(SPIN_CYCLES)
        a.record(torch.cuda.current_stream())
        g.replay()
        b.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) / reps)
    return float(np.median(times))


def eager_time(torch, fn, reps, spin=SPIN_CYCLES, refresh=None):
    """ms per fn(i) over reps eager calls on the current stream.  As in
    graph_time, a spin of `spin` GPU cycles is queued ahead of the start
    event, so that the region holds the calls' GPU work back to back and not
    the host's latency before the first launch (callers whose Python-side
    call is slow pass a longer spin); refresh() runs first, if given."""
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    if refresh:
        refresh()
    torch.cuda.synchronize()
    torch.cuda._sleep(spin)
    a.record()
    for i in range(reps):
        fn(i)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


_READ_FLUSH = []


def read_flush(torch):
    """Reads 512 MiB of a buffer of its own (a reduction), so that what a
    refresh has just written leaves the 256 MB MALL, written back, before
    the timed region instead of inside it."""
    if not _READ_FLUSH:
        _READ_FLUSH.append(torch.ones(128 << 20, dtype=torch.float32, device="cuda"))
    _READ_FLUSH[0].sum()


class Rotation(list):
    """Buffers of one shape whose total exceeds ROTATE_BYTES, and at least
    `count` of them, each = init.  Transforms run in place; repeated on one
    buffer they grow fp16 values to inf/NaN within two or three transforms,
    and such data runs faster than finite data (a few percent on this part).
    refresh() puts init back into every buffer (from a device copy), then
    runs read_flush: a timed region with at most one transform per buffer
    then sees finite data, cold in the MALL."""

    def __init__(self, torch, shape, dtype, init, count=0):
        one = int(np.prod(shape)) * torch.tensor([], dtype=dtype).element_size()
        k = max(2, count, -(-ROTATE_BYTES // one))
        self.torch = torch
        self.src = init.to(device="cuda", dtype=dtype).reshape(shape)
        super().__init__(torch.empty(shape, dtype=dtype, device="cuda") for _ in range(k))
        for b in self:
            b.copy_(self.src)

    def refresh(self):
        for b in self:
            b.copy_(self.src)
        read_flush(self.torch)


def rotation(torch, shape, dtype, init, count=0):
    """Rotation(torch, shape, dtype, init, count)."""
    return Rotation(torch, shape, dtype, init, count)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def ref_error(tested, standard):
    """The reference harness's error statistic (accuracy.cpp:25-41): per real
    component min(|(t - s) / s|, 1), summed and divided by n * batch."""
    t = np.stack([tested.real, tested.imag], -1).ravel()
    s = np.stack([standard.real, standard.imag], -1).ravel()
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.abs((t - s) / s)
    return float(np.where(v < 1.0, v, 1.0).sum() / (tested.size))


def to_c(t):
    v = t.float().cpu().numpy().astype(np.float64)
    return v[..., 0] + 1j * v[..., 1]


def fp16_uniform(rng, shape):
    return rng.uniform(-1, 1, tuple(shape) + (2,)).astype(np.float16)


def summary(ms, n_points, batch, alg_bytes):
    flops = 5.0 * n_points * math.log2(n_points) * batch
    return {"ms": round(ms, 5), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
            "hbm_gbs_alg": round(alg_bytes / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(alg_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}


# ----------------------------------------------------------------------------
# CPU baselines

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
        if el > 10.0 or reps >= 5000:
            break
    flops = 5.0 * n * math.log2(n) * sig * reps
    return {"value": round(flops / el / 1e9, 3), "unit": "GFLOP/s", "cores": 1, "kind": "port",
            "sample": f"{reps} x {sig} signals of N={n} (FP64 DIT restatement, {el:.1f} s)"}


def cpu_threads() -> int:
    """Host threads of this process's CPU share (affinity, capped by OMP_NUM_THREADS)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# FFTW3-API libraries for the single-precision host FFT, in order of preference:
# FFTW itself, then MKL's FFTW3 wrappers; scipy.fft is the last resort.
FFTW_LIBS = ("libfftw3f.so.3", "libfftw3f.so", "/opt/conda/lib/libmkl_rt.so")


def fftw_child(n: int, sig: int, threads: int, seconds: float):
    """Runs in a child process (the FFTW3 library's threading runtime never
    meets torch's): fftwf_plan_many_dft over `sig` signals of n, in place,
    complex64, FFTW_MEASURE, `threads` threads; one JSON line on stdout."""
    import ctypes
    lib, name = None, None
    for cand in FFTW_LIBS:
        try:
            L = ctypes.CDLL(cand)
            getattr(L, "fftwf_plan_many_dft")
            lib, name = L, cand
            break
        except (OSError, AttributeError):
            continue
    if lib is None:
        print(json.dumps({"error": "no FFTW3-API library"}))
        return
    vp, ip = ctypes.c_void_p, ctypes.c_int
    lib.fftwf_plan_many_dft.restype = vp
    lib.fftwf_plan_many_dft.argtypes = [ip, ctypes.POINTER(ip), ip, vp, vp, ip, ip, vp, vp, ip, ip, ip, ctypes.c_uint]
    lib.fftwf_execute.argtypes = [vp]
    lib.fftwf_destroy_plan.argtypes = [vp]
    if hasattr(lib, "fftwf_init_threads"):
        lib.fftwf_init_threads()
        lib.fftwf_plan_with_nthreads.argtypes = [ip]
        lib.fftwf_plan_with_nthreads(threads)
    rng = np.random.default_rng(0)
    x = (rng.uniform(-1, 1, (sig, n)) + 1j * rng.uniform(-1, 1, (sig, n))).astype(np.complex64)
    buf = np.empty_like(x)
    dims = (ip * 1)(n)
    plan = lib.fftwf_plan_many_dft(1, dims, sig, buf.ctypes.data, None, 1, n, buf.ctypes.data, None, 1, n, -1, 0)
    if not plan:
        print(json.dumps({"error": f"{name}: planning failed"}))
        return
    buf[:] = x
    lib.fftwf_execute(plan)
    err = float(np.linalg.norm(buf[0] - np.fft.fft(x[0].astype(np.complex128)))
                / np.linalg.norm(np.fft.fft(x[0].astype(np.complex128))))
    reps = 0
    t0 = time.perf_counter()
    while True:
        lib.fftwf_execute(plan)
        reps += 1
        el = time.perf_counter() - t0
        if el > seconds or reps >= 100000:
            break
    lib.fftwf_destroy_plan(plan)
    print(json.dumps({"library": name, "reps": reps, "seconds": el, "rel_l2_first_signal": err}))


def cpu_fft_f32(n: int, batch: int):
    """The single-precision host FFT on this process's CPU share, bounded
    sample of the workload: an FFTW3-API library (FFTW, else MKL's wrappers)
    in a child process, else scipy.fft (pocketfft)."""
    cores = cpu_threads()
    sig = min(batch, 64)
    flops_per = 5.0 * n * math.log2(n) * sig
    res = None
    try:
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), MKL_NUM_THREADS=str(cores))
        cp = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-fft-child", str(n), str(sig),
                             str(cores)], capture_output=True, text=True, timeout=120, env=env)
        lines = [l for l in cp.stdout.splitlines() if l.startswith("{")]
        if cp.returncode == 0 and lines:
            r = json.loads(lines[-1])
            if "error" not in r and r["rel_l2_first_signal"] < 1e-5:
                res = r
    except (subprocess.SubprocessError, OSError, ValueError, KeyError):
        res = None
    if res is not None:
        reps, el, library = res["reps"], res["seconds"], f"{res['library']} (fftwf_plan_many_dft, complex64, in place)"
    else:
        import scipy.fft
        rng = np.random.default_rng(0)
        x = (rng.uniform(-1, 1, (sig, n)) + 1j * rng.uniform(-1, 1, (sig, n))).astype(np.complex64)
        scipy.fft.fft(x, axis=-1, workers=cores)  # warm-up (plan, threads)
        reps = 0
        t0 = time.perf_counter()
        while True:
            scipy.fft.fft(x, axis=-1, workers=cores, overwrite_x=False)
            reps += 1
            el = time.perf_counter() - t0
            if el > 5.0 or reps >= 200:
                break
        library = f"scipy.fft {scipy.__version__} (pocketfft, complex64; no FFTW3-API library loaded)"
    gf = flops_per * reps / el / 1e9
    return {"value": round(gf, 2), "unit": "GFLOP/s", "cores": cores, "cpu_model": cpu_model(), "library": library,
            "sample": f"{reps} x {sig} signals of N={n} ({el:.1f} s)",
            "ms_full_batch_est": round(5.0 * n * math.log2(n) * batch / (gf * 1e9) * 1e3, 3)}


# ----------------------------------------------------------------------------
# workloads

def bench_plan1d(torch, tcfft_amd, n, batch, reps, rng):
    plan = tcfft_amd.Plan1D(n, batch)
    x = fp16_uniform(rng, (batch, n))
    xr = np.ascontiguousarray(x[:, tcfft_amd.gen_rev(n, plan.radices)])
    bufs = rotation(torch, (batch, n, 2), torch.float16, torch.from_numpy(xr), count=reps)
    sched = plan.autotune(bufs[0].data_ptr())[:2]  # start schedule on this box (cold data)
    bufs[0].copy_(torch.from_numpy(xr))
    plan.exec(bufs[0].data_ptr())
    torch.cuda.synchronize()
    k = min(batch, 4)
    err = rel(to_c(bufs[0][:k]), np.fft.fft(x[:k, :, 0].astype(np.float64) + 1j * x[:k, :, 1], axis=-1))
    ms = graph_time(torch, lambda i, st: plan.exec(bufs[i % len(bufs)].data_ptr(), tcfft_amd.FORWARD, st), reps, 3,
                    refresh=bufs.refresh)
    out = summary(ms, n, batch, 8.0 * n * batch)
    out.update({"rel_l2_vs_fp64": err, "passes": plan.num_passes, "radices": plan.radices,
                "start_schedule": list(sched)})
    # natural-order input (tcfft_exec_natural, out of place) on the same data
    nat = rotation(torch, (batch, n, 2), torch.float16, torch.from_numpy(np.ascontiguousarray(x)))
    dst = torch.empty_like(nat[0])
    out["ms_natural_order"] = round(graph_time(
        torch, lambda i, st: plan.exec_natural(nat[i % len(nat)].data_ptr(), dst.data_ptr(), tcfft_amd.FORWARD, st),
        reps), 5)
    del nat
    plan.destroy()
    return out


def bench_plan2d(torch, tcfft_amd, nx, ny, batch, reps, rng):
    plan = tcfft_amd.Plan2D(nx, ny, batch)
    x = fp16_uniform(rng, (batch, nx, ny))
    rx = tcfft_amd.gen_rev(nx, plan.radices_x)
    ry = tcfft_amd.gen_rev(ny, plan.radices_y)
    xr = np.ascontiguousarray(x[:, rx][:, :, ry])
    bufs = rotation(torch, (batch, nx, ny, 2), torch.float16, torch.from_numpy(xr), count=reps)
    sched = plan.autotune(bufs[0].data_ptr())[:2]  # start schedules on this box (cold data)
    bufs[0].copy_(torch.from_numpy(xr))
    plan.exec(bufs[0].data_ptr())
    torch.cuda.synchronize()
    err = rel(to_c(bufs[0][:1]), np.fft.fft2(x[:1, :, :, 0].astype(np.float64) + 1j * x[:1, :, :, 1]))
    ms = graph_time(torch, lambda i, st: plan.exec(bufs[i % len(bufs)].data_ptr(), tcfft_amd.FORWARD, st), reps, 3,
                    refresh=bufs.refresh)
    out = summary(ms, nx * ny, batch, 8.0 * nx * ny * batch)
    out.update({"rel_l2_vs_fp64": err, "passes": plan.num_passes, "radices_x": plan.radices_x,
                "radices_y": plan.radices_y, "start_schedule": list(sched)})
    nat = rotation(torch, (batch, nx, ny, 2), torch.float16, torch.from_numpy(np.ascontiguousarray(x)))
    dst = torch.empty_like(nat[0])
    out["ms_natural_order"] = round(graph_time(
        torch, lambda i, st: plan.exec_natural(nat[i % len(nat)].data_ptr(), dst.data_ptr(), tcfft_amd.FORWARD, st),
        reps), 5)
    del nat
    plan.destroy()
    return out


def bench_split(torch, tcfft_amd, reps, chunk=None, ms_plain=None):
    """Split precision, N = 2^16 x 256, FP32 complex in/out,
    N(0,1)*10^U(-3,3) per signal; time ratio to the plain fp16 exec."""
    n, batch = 1 << 16, 256
    rng = np.random.default_rng(0)
    scale = 10.0 ** rng.uniform(-3, 3, (batch, 1))
    x = ((rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))) * scale).astype(np.complex64)
    host = torch.from_numpy(x.view(np.float32))
    ins = rotation(torch, tuple(host.shape), torch.float32, host)
    outs = [torch.empty_like(ins[0]) for _ in ins]
    plan = tcfft_amd.Plan1D(n, batch)
    if chunk is not None:
        plan.set_option(tcfft_amd.OPT_SPLIT_CHUNK, chunk)
    plan.exec_split(ins[0].data_ptr(), outs[0].data_ptr())
    torch.cuda.synchronize()
    k = 8
    got = outs[0][:k].cpu().numpy().view(np.complex64).astype(np.complex128)
    ref = np.fft.fft(x[:k].astype(np.complex128), axis=-1)
    errs = [rel(got[i], ref[i]) for i in range(k)]
    ms = graph_time(torch, lambda i, st: plan.exec_split(ins[i % len(ins)].data_ptr(), outs[i % len(ins)].data_ptr(),
                                                         tcfft_amd.FORWARD, st), reps, 3)
    if ms_plain is None:
        pb = rotation(torch, (batch, n, 2), torch.float16,
                      torch.from_numpy(fp16_uniform(np.random.default_rng(1), (batch, n))), count=reps)
        ms_plain = graph_time(torch, lambda i, st: plan.exec(pb[i % len(pb)].data_ptr(), tcfft_amd.FORWARD, st), reps, 3,
                              refresh=pb.refresh)
    plan.destroy()
    out = summary(ms, n, batch, 16.0 * n * batch)
    out.update({"rel_l2_max_per_signal": max(errs), "ms_plain_fp16": round(ms_plain, 5),
                "time_ratio_vs_plain": round(ms / ms_plain, 3), "split_chunk": chunk})
    return out


def bench_split_shapes(torch, tcfft_amd, reps):
    """Split precision on prepared plans at the other configured shapes
    (1D 2^20 x 16, 2D 1024^2 x 8), FP32 in/out out of place, rotated
    buffers; rel-L2 of the first signal / image against FP64."""
    res = {}
    for name, dims, a, b, batch in (("split_n2^20_b16", 1, 1 << 20, 0, 16),
                                    ("split_2d_1024x1024_b8", 2, 1024, 1024, 8)):
        plan = tcfft_amd.Plan1D(a, batch) if dims == 1 else tcfft_amd.Plan2D(a, b, batch)
        plan.prepare_split()
        per = a * (b or 1)
        rng = np.random.default_rng(4)
        x = (rng.standard_normal((batch, per)) + 1j * rng.standard_normal((batch, per))).astype(np.complex64)
        host = torch.from_numpy(x.view(np.float32))
        ins = rotation(torch, tuple(host.shape), torch.float32, host)
        outs = [torch.empty_like(ins[0]) for _ in ins]
        plan.exec_split(ins[0].data_ptr(), outs[0].data_ptr())
        torch.cuda.synchronize()
        got = outs[0][0].cpu().numpy().view(np.complex64).astype(np.complex128)
        x0 = x[0].astype(np.complex128)
        ref = np.fft.fft(x0) if dims == 1 else np.fft.fft2(x0.reshape(a, b)).reshape(-1)
        err = rel(got, ref)
        ms = graph_time(torch, lambda i, st: plan.exec_split(ins[i % len(ins)].data_ptr(),
                                                             outs[i % len(ins)].data_ptr(), tcfft_amd.FORWARD, st),
                        reps, 3)
        npass = len(tcfft_amd.query_split(dims, a, b, batch))
        res[name] = {"ms": round(ms, 5), "passes": npass, "rel_l2_vs_fp64": err,
                     "hbm_frac_per_pass": round(16.0 * per * batch * npass / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
        plan.destroy()
        del ins, outs
    return res


def bench_uneven(torch, tcfft_amd, reps):
    """4096 fp16 signals, N = 2^U{8..16} (seed 0), drop-in order: one
    persistent autotuned launch vs one launch per size class (regular plans
    on a size-sorted layout of the same volume)."""
    rng = np.random.default_rng(0)
    n_sig = 4096
    sizes = (2 ** rng.integers(8, 17, n_sig)).astype(np.int32)
    total = int(sizes.sum())
    flat = rng.uniform(-1, 1, (total, 2)).astype(np.float16)
    mp = tcfft_amd.MultiPlan(sizes)
    host = torch.from_numpy(flat)
    bufs = rotation(torch, tuple(host.shape), torch.float16, host, count=reps)
    g, o, _ = mp.autotune(bufs[0].data_ptr())
    delay = mp.schedule()[2]
    ms_multi = graph_time(torch, lambda i, st: mp.exec(bufs[i % len(bufs)].data_ptr(), tcfft_amd.FORWARD, st), reps, 3,
                          refresh=bufs.refresh)
    classes = sorted(set(int(v) for v in sizes))
    plans, offs, o2 = [], [], 0
    for n in classes:
        cnt = int((sizes == n).sum())
        plans.append(tcfft_amd.Plan1D(n, cnt))
        offs.append(o2)
        o2 += n * cnt

    def per_size(i, st):
        base = bufs[i % len(bufs)].data_ptr()
        for p, off in zip(plans, offs):
            p.exec(base + 4 * off, tcfft_amd.FORWARD, st)

    ms_sizes = graph_time(torch, per_size, reps, 3, refresh=bufs.refresh)
    # natural-order signals (tcfft_multi_exec_natural, out of place, the same schedule)
    nat_out = torch.empty_like(bufs[0])
    ms_nat = graph_time(torch, lambda i, st: mp.exec_natural(bufs[i % len(bufs)].data_ptr(), nat_out.data_ptr(),
                                                               tcfft_amd.FORWARD, st), reps, 3, refresh=bufs.refresh)
    # accuracy on a sample of signals, digit-reversed per signal
    sample = [int(i) for i in rng.choice(n_sig, 16, replace=False)]
    check = torch.from_numpy(flat).cuda()
    for i in sample:
        n, o3 = int(sizes[i]), int(mp.offsets[i])
        rad, _ = tcfft_amd.query_1d(n, 1)
        check[o3:o3 + n] = torch.from_numpy(np.ascontiguousarray(flat[o3:o3 + n][tcfft_amd.gen_rev(n, rad)])).cuda()
    mp.exec(check.data_ptr())
    torch.cuda.synchronize()
    errs = []
    for i in sample:
        n, o3 = int(sizes[i]), int(mp.offsets[i])
        x = flat[o3:o3 + n].astype(np.float64)
        errs.append(rel(to_c(check[o3:o3 + n]), np.fft.fft(x[:, 0] + 1j * x[:, 1])))
    flops = float(sum(5.0 * n * math.log2(n) for n in sizes))
    n_tasks = mp.n_tasks
    for p in plans:
        p.destroy()
    mp.destroy()
    return {"ms": round(ms_multi, 5), "gflops": round(flops / (ms_multi * 1e-3) / 1e9, 1),
            "hbm_frac": round(8.0 * total / (ms_multi * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "ms_per_size_launches": round(ms_sizes, 5), "speedup_vs_per_size": round(ms_sizes / ms_multi, 3),
            "ms_natural_order": round(ms_nat, 5),
            "signals": n_sig, "elements": total, "tasks": n_tasks, "autotuned_grid": g, "autotuned_order": o, "autotuned_start_delay": delay,
            "rel_l2_max_sample": max(errs)}


def bench_conv(torch, reps):
    """FFT-conv layer y = F^H (H * F x) / n, x [64, 256, 8192] real-as-complex
    fp16, filter spectrum [256, 8192] complex ~ N(0, 0.02): forward and
    forward + backward per step (eager, torch autograd over the fused HIP
    convolution kernels)."""
    from tcfft_amd import torch_fft
    torch.manual_seed(0)
    B, C, n = 64, 256, 8192
    layer = torch_fft.FFTConv1d(C, n, std=0.02, device="cuda")
    xr = (torch.rand(B, C, n, device="cuda") * 2 - 1).half()
    x = torch.stack([xr, torch.zeros_like(xr)], dim=-1).requires_grad_(True)
    with torch.no_grad():
        y = layer(x)
    torch.cuda.synchronize()
    hc = layer.weight.detach().double().cpu().numpy()
    hc = hc[..., 0] + 1j * hc[..., 1]
    errs = []
    for (b, c) in ((0, 0), (63, 255)):
        xs = x[b, c].detach().double().cpu().numpy()
        ref = np.fft.ifft(np.fft.fft(xs[:, 0] + 1j * xs[:, 1]) * hc[c])
        errs.append(rel(to_c(y[b, c]), ref))

    def fwd(i):
        with torch.no_grad():
            layer(x)

    g = torch.randn(B, C, n, 2, device="cuda").half()

    def fwd_bwd(i):
        x.grad = None
        layer.weight.grad = None
        layer(x).backward(g)

    fwd(0)
    fwd_bwd(0)
    # a ~2 ms spin: the layer's Python forward and backward take far longer
    # on the host than the default spin covers
    ms_f = eager_time(torch, fwd, reps, spin=5_000_000)
    ms_fb = eager_time(torch, fwd_bwd, max(2, reps // 2), spin=5_000_000)
    sig = B * C
    fft_flops = 5.0 * n * math.log2(n) * sig
    return {"ms_forward": round(ms_f, 4), "ms_forward_backward": round(ms_fb, 4),
            "fft_gflops_forward": round(2 * fft_flops / (ms_f * 1e-3) / 1e9, 1),
            "signals": sig, "n": n, "rel_l2_forward_sample": max(errs),
            "note": "forward = one fused kernel launch (k_conv: natural input -> DIF -> filter product in "
                    "registers -> inverse DIT -> output) after the filter's digit-reversal permutation; "
                    "forward+backward adds the conjugate-filter launch for the input gradient and the "
                    "filter-gradient reduction over the stored spectra (torch autograd)"}


def bench_r2c(torch, tcfft_amd, reps):
    """Real-input transforms on the FFT-conv input shape as real signals:
    16384 = 64 x 256 signals of n = 8192 fp16 reals, forward (bins 0..n/2)
    and inverse, beside the complex transform of the same data fed as complex
    (exec_natural, out of place), which is what a caller without R2C runs."""
    n, batch = 8192, 64 * 256
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (batch, n)).astype(np.float16)
    plan = tcfft_amd.PlanR2C(n, batch)
    bins = n // 2 + 1
    real = rotation(torch, (batch, n), torch.float16, torch.from_numpy(x))
    spec = [torch.empty(batch, bins, 2, dtype=torch.float16, device="cuda") for _ in range(3)]
    plan.exec_r2c(real[0].data_ptr(), spec[0].data_ptr())
    torch.cuda.synchronize()
    k = 4
    err_f = rel(to_c(spec[0][:k]), np.fft.rfft(x[:k].astype(np.float64), axis=-1))
    for sp in spec[1:]:
        sp.copy_(spec[0])
    back = torch.empty(batch, n, dtype=torch.float16, device="cuda")
    plan.exec_c2r(spec[0].data_ptr(), back.data_ptr())
    torch.cuda.synchronize()
    ref_b = n * np.fft.irfft(to_c(spec[0][:k]), n, axis=-1)
    err_i = rel(back[:k].double().cpu().numpy(), ref_b)
    ms_f = graph_time(torch, lambda i, st: plan.exec_r2c(real[i % len(real)].data_ptr(), spec[i % 3].data_ptr(), st),
                      reps, 3)
    ms_i = graph_time(torch, lambda i, st: plan.exec_c2r(spec[i % 3].data_ptr(), real[i % len(real)].data_ptr(), st),
                      reps, 3)
    plan.destroy()
    del real, spec, back
    alg = (2.0 * n + 4.0 * bins) * batch
    cp = tcfft_amd.Plan1D(n, batch)
    xc = np.zeros((batch, n, 2), dtype=np.float16)
    xc[..., 0] = x
    src = rotation(torch, (batch, n, 2), torch.float16, torch.from_numpy(xc))
    dst = torch.empty_like(src[0])
    ms_c = graph_time(torch, lambda i, st: cp.exec_natural(src[i % len(src)].data_ptr(), dst.data_ptr(),
                                                           tcfft_amd.FORWARD, st), reps, 3)
    cp.destroy()
    del src, dst
    fwd = summary(ms_f, n, batch, alg)
    return {"n": n, "signals": batch, "ms_r2c": round(ms_f, 5), "ms_c2r": round(ms_i, 5),
            "hbm_frac_r2c": fwd["hbm_frac"], "hbm_frac_c2r": summary(ms_i, n, batch, alg)["hbm_frac"],
            "alg_bytes": int(alg), "rel_l2_r2c": err_f, "rel_l2_c2r": err_i,
            "ms_complex_of_real_data": round(ms_c, 5), "speedup_vs_complex": round(ms_c / ms_f, 3)}


def hipfft_fp16(torch, dims, batch, reps, fp32=False):
    """The vendor library on the same workload: hipFFT (rocFFT) half-precision
    C2C through hipfftXtMakePlanMany(HIP_C_16F), in place, natural order,
    timed with events over `reps` eager execs after a warm-up.  fp32: the
    single-precision C2C (HIP_C_32F) instead, the vendor counterpart of the
    split-precision path."""
    import ctypes
    try:
        lib = ctypes.CDLL("/opt/rocm/lib/libhipfft.so")
    except OSError as e:
        return {"error": f"libhipfft not loadable: {e}"}
    ctype = 4 if fp32 else 6  # HIP_C_32F / HIP_C_16F
    h = ctypes.c_void_p()
    if lib.hipfftCreate(ctypes.byref(h)) != 0:
        return {"error": "hipfftCreate failed"}
    rank = len(dims)
    n = (ctypes.c_longlong * rank)(*dims)
    total = int(np.prod(dims))
    ws = ctypes.c_size_t()
    rc = lib.hipfftXtMakePlanMany(h, rank, n, None, ctypes.c_longlong(1), ctypes.c_longlong(total), ctype, None,
                                  ctypes.c_longlong(1), ctypes.c_longlong(total), ctype, ctypes.c_longlong(batch),
                                  ctypes.byref(ws), ctype)
    if rc != 0:
        lib.hipfftDestroy(h)
        return {"error": f"hipfftXtMakePlanMany returned {rc}"}
    stream = torch.cuda.current_stream().cuda_stream
    lib.hipfftSetStream(h, ctypes.c_void_p(stream))
    rng = np.random.default_rng(2)
    x = fp16_uniform(rng, (batch,) + tuple(dims))
    if fp32:
        x = x.astype(np.float32)
    bufs = rotation(torch, x.shape, torch.float32 if fp32 else torch.float16, torch.from_numpy(x), count=reps)
    lib.hipfftXtExec(h, ctypes.c_void_p(bufs[0].data_ptr()), ctypes.c_void_p(bufs[0].data_ptr()), -1)
    torch.cuda.synchronize()
    k = min(batch, 2)
    xs = x[:k, ..., 0].astype(np.float64) + 1j * x[:k, ..., 1]
    ref = np.fft.fftn(xs, axes=tuple(range(1, rank + 1)))
    err = rel(to_c(bufs[0][:k]), ref)
    nb = len(bufs)
    ms = eager_time(torch, lambda i: lib.hipfftXtExec(h, ctypes.c_void_p(bufs[i % nb].data_ptr()),
                                                       ctypes.c_void_p(bufs[i % nb].data_ptr()), -1), reps,
                    refresh=bufs.refresh)
    lib.hipfftDestroy(h)
    out = summary(ms, total, batch, (16.0 if fp32 else 8.0) * total * batch)
    out.update({"rel_l2_vs_fp64": err,
                "library": f"hipFFT/rocFFT (/opt/rocm), {'HIP_C_32F' if fp32 else 'HIP_C_16F'}, in place, eager",
                "work_bytes": int(ws.value)})
    return out


def hipfft_r2c_fp16(torch, n, batch, reps):
    """hipFFT's half-precision R2C on the same real workload
    (hipfftXtMakePlanMany, HIP_R_16F -> HIP_C_16F, out of place), eager."""
    import ctypes
    try:
        lib = ctypes.CDLL("/opt/rocm/lib/libhipfft.so")
    except OSError as e:
        return {"error": f"libhipfft not loadable: {e}"}
    HIP_R_16F, HIP_C_16F = 2, 6
    h = ctypes.c_void_p()
    if lib.hipfftCreate(ctypes.byref(h)) != 0:
        return {"error": "hipfftCreate failed"}
    dims = (ctypes.c_longlong * 1)(n)
    bins = n // 2 + 1
    ws = ctypes.c_size_t()
    rc = lib.hipfftXtMakePlanMany(h, 1, dims, None, ctypes.c_longlong(1), ctypes.c_longlong(n), HIP_R_16F, None,
                                  ctypes.c_longlong(1), ctypes.c_longlong(bins), HIP_C_16F, ctypes.c_longlong(batch),
                                  ctypes.byref(ws), HIP_C_16F)
    if rc != 0:
        lib.hipfftDestroy(h)
        return {"error": f"hipfftXtMakePlanMany (R2C) returned {rc}"}
    lib.hipfftSetStream(h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (batch, n)).astype(np.float16)
    src = rotation(torch, (batch, n), torch.float16, torch.from_numpy(x))
    dst = torch.empty(batch, bins, 2, dtype=torch.float16, device="cuda")
    lib.hipfftXtExec(h, ctypes.c_void_p(src[0].data_ptr()), ctypes.c_void_p(dst.data_ptr()), -1)
    torch.cuda.synchronize()
    k = min(batch, 2)
    err = rel(to_c(dst[:k]), np.fft.rfft(x[:k].astype(np.float64), axis=-1))
    ns = len(src)
    ms = eager_time(torch, lambda i: lib.hipfftXtExec(h, ctypes.c_void_p(src[i % ns].data_ptr()),
                                                       ctypes.c_void_p(dst.data_ptr()), -1), reps)
    lib.hipfftDestroy(h)
    return {"ms": round(ms, 5), "rel_l2_vs_fp64": err, "work_bytes": int(ws.value),
            "library": "hipFFT/rocFFT (/opt/rocm), HIP_R_16F -> HIP_C_16F, out of place, eager"}


def stream_copy(torch):
    """Device-to-device copy bandwidth over a > MALL rotation (TB/s, read+write)."""
    src = rotation(torch, (64 << 20,), torch.float32, torch.zeros(64 << 20))  # 256 MiB each
    dst = [torch.empty_like(src[0]) for _ in range(2)]
    ms = eager_time(torch, lambda i: dst[i % 2].copy_(src[i % len(src)]), 20)
    return round(2 * src[0].numel() * 4 / (ms * 1e-3) / 1e12, 3)


# ----------------------------------------------------------------------------

def load_traffic(n: int, batch: int):
    """Counter-derived HBM bytes per launch of the 65536-point kernel, if profiled."""
    path = os.path.join(ROOT, "profiles", f"traffic_n{n}_b{batch}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def main():
    if len(sys.argv) == 5 and sys.argv[1] == "--cpu-fft-child":
        fftw_child(int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4]), 5.0)
        return
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): every workload here shards on "
              f"the batch axis with no exchange, and this build runs on one GPU. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import tcfft_amd

    torch.cuda.set_device(0)
    if args.workload != "plain":
        leg = {"split": lambda: bench_split(torch, tcfft_amd, args.steps, args.split_chunk),
               "uneven": lambda: bench_uneven(torch, tcfft_amd, args.steps),
               "conv": lambda: bench_conv(torch, max(3, args.steps // 20)),
               "r2c": lambda: {**bench_r2c(torch, tcfft_amd, args.steps),
                               "hipfft_r2c": hipfft_r2c_fp16(torch, 8192, 64 * 256, args.steps)}}[args.workload]()
        print(json.dumps({"workload": args.workload, **leg}), flush=True)
        return

    n, batch = args.n, args.batch
    plan = tcfft_amd.Plan1D(n, batch)
    rng = np.random.default_rng(0)
    x = fp16_uniform(rng, (batch, n))
    x_rev = np.ascontiguousarray(x[:, tcfft_amd.gen_rev(n, plan.radices)])
    host = torch.from_numpy(x_rev)
    # one transform per buffer per timed region (per-pass timing: 100 launches)
    bufs = rotation(torch, (batch, n, 2), torch.float16, host, count=max(args.steps, 100))
    torch.cuda.synchronize()
    if args.calib:
        for _ in range(5):
            bufs[1].copy_(bufs[0])
        torch.cuda.synchronize()

    # the project's own start-schedule autotuner picks the first pass's
    # schedule on this box (cold data: a 256 MiB write before each timed run)
    tuned_groups, tuned_delay, tuned_ms = plan.autotune(bufs[0].data_ptr())
    bufs[0].copy_(host)
    torch.cuda.synchronize()

    # accuracy of one exec vs FP64 (numpy pocketfft) on a few signals
    plan.exec(bufs[0].data_ptr())
    torch.cuda.synchronize()
    k = min(batch, 8)
    ref64 = np.fft.fft(x[:k, :, 0].astype(np.float64) + 1j * x[:k, :, 1], axis=-1)
    rel_l2 = rel(to_c(bufs[0][:k]), ref64)
    ref_err = ref_error(to_c(bufs[0][:k]), ref64)
    bufs[0].copy_(host)

    nb = len(bufs)
    for i in range(args.warmup):
        plan.exec(bufs[i % nb].data_ptr())
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    ms_eager = eager_time(torch, lambda i: plan.exec(bufs[i % nb].data_ptr()), args.steps, refresh=bufs.refresh)
    ms = graph_time(torch, lambda i, st: plan.exec(bufs[i % nb].data_ptr(), tcfft_amd.FORWARD, st), args.steps,
                    refresh=bufs.refresh)
    t_wall = time.perf_counter() - t_wall0

    # per-pass timing (each pass kernel alone, graph-replayed on its stream)
    per_pass = []
    for p in range(plan.num_passes):
        per_pass.append(graph_time(torch, lambda i, st, p=p: plan.exec_passes(bufs[i % nb].data_ptr(), p, 1,
                                                                              tcfft_amd.FORWARD, st),
                                   max(100, args.steps), 5, refresh=bufs.refresh))
    dom = int(np.argmax(per_pass))
    alg_bytes = 8.0 * n * batch
    achieved = alg_bytes / (per_pass[dom] * 1e-3) / 1e9

    # natural-order input (tcfft_exec_natural, out of place) on the same
    # signals, under the natural kernel's own default start schedule
    del bufs
    plan.set_option(tcfft_amd.OPT_BIG_DELAY, -1)
    nat = rotation(torch, (batch, n, 2), torch.float16, torch.from_numpy(np.ascontiguousarray(x)))
    dst = torch.empty_like(nat[0])
    plan.exec_natural(nat[0].data_ptr(), dst.data_ptr())
    torch.cuda.synchronize()
    nat_rel_l2 = rel(to_c(dst[:k]), ref64)
    ms_nat = graph_time(torch, lambda i, st: plan.exec_natural(nat[i % len(nat)].data_ptr(), dst.data_ptr(),
                                                               tcfft_amd.FORWARD, st), args.steps)
    del nat, dst
    flops = 5.0 * n * math.log2(n) * batch
    out = {
        "metric": METRIC,
        "value": round(flops / (ms * 1e-3) / 1e9, 1),
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
        "data": f"synthetic uniform[-1,1) fp16, resident in HBM, rotated over {nb} buffers, each transformed "
                "at most once per timed region (input restored before each)",
        "config": {"workload": f"1D C2C FFT N={n} batch={batch}, fp16 interleaved, in place, "
                               f"digit-reversed input (tcfftExec)", "n": n, "batch": batch,
                   "passes": plan.num_passes, "radices": plan.radices,
                   "start_schedule": {"groups": tuned_groups, "delay": tuned_delay,
                                      "tuner_ms": round(tuned_ms, 5),
                                      "picked_by": "tcfft_plan_autotune on this box (cold data)"}},
        "rel_l2_vs_fp64": rel_l2,
        "ref_error_metric": ref_err,
        "hbm_frac_whole_exec": round(alg_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        "pass_ms": [round(v, 5) for v in per_pass],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": load_traffic(n, batch),
                     "kernel": f"pass {dom} of {plan.num_passes}"},
        "wall_s_timed": round(t_wall, 3),
        "natural_order": {"ms": round(ms_nat, 5), "rel_l2_vs_fp64": nat_rel_l2,
                          "hbm_frac": round(alg_bytes / (ms_nat * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                          "note": "tcfft_exec_natural, natural-order input, out of place, same signals"},
    }
    plan.destroy()
    if not args.no_configs and n == N and batch == BATCH:
        reps = max(20, args.steps // 4)
        crng = np.random.default_rng(1)
        cfg = {}
        cfg["c1_n4096_b1024"] = bench_plan1d(torch, tcfft_amd, 4096, 1024, reps, crng)
        cfg["c2a_n2^20_b16"] = bench_plan1d(torch, tcfft_amd, 1 << 20, 16, reps, crng)
        cfg["c2b_2d_1024x1024_b8"] = bench_plan2d(torch, tcfft_amd, 1024, 1024, 8, reps, crng)
        cfg["c3_split_n2^16_b256"] = bench_split(torch, tcfft_amd, reps, ms_plain=ms)
        cfg.update(bench_split_shapes(torch, tcfft_amd, reps))
        cfg["c4_uneven_4096_signals"] = bench_uneven(torch, tcfft_amd, reps)
        cfg["c5_fftconv_64x256x8192"] = bench_conv(torch, 10)
        cfg["r2c_n8192_16384_signals"] = bench_r2c(torch, tcfft_amd, reps)
        cfg["hipfft_fp16_n2^16_b256"] = hipfft_fp16(torch, [n], batch, reps)
        cfg["hipfft_fp16_n4096_b1024"] = hipfft_fp16(torch, [4096], 1024, reps)
        cfg["hipfft_fp16_2d_1024x1024_b8"] = hipfft_fp16(torch, [1024, 1024], 8, reps)
        cfg["hipfft_fp16_n2^20_b16"] = hipfft_fp16(torch, [1 << 20], 16, reps)
        cfg["hipfft_fp32_n2^16_b256"] = hipfft_fp16(torch, [n], batch, reps, fp32=True)
        cfg["hipfft_r2c_fp16_n8192_16384_signals"] = hipfft_r2c_fp16(torch, 8192, 64 * 256, reps)
        cfg["stream_copy_tbs"] = stream_copy(torch)
        out["configs"] = cfg
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n)
        out["cpu_fft_f32"] = cpu_fft_f32(n, batch)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
