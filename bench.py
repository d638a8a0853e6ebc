#!/usr/bin/env python3
"""Headline benchmark: batched 1024^2 angular-spectrum propagation on one GPU.

Workload (one "step"): asm_propagate of a batch of 64 complex64 1024x1024
planes, each exp(i phi) with phi ~ U[0, 2 pi) from numpy default_rng(0),
dx = 0.3 um, lambda = 0.532 um, n = 1.33, z = 100 um, pad_width = 0,
mode = "same" (the headline configuration of the propagation path; 64
planes = 512 MiB, twice the Infinity Cache).  Inputs are resident in HBM
before the timed region.

Prints ONE JSON line on stdout.  ``value`` is propagations/s over the timed
steps.  ``roofline`` describes the dominant kernel, timed live with HIP
events on its launch stream (libcxm's timing hooks); ``path_roofline`` the
whole propagation under the algorithmic cost model of 32 N^2 bytes per
propagation (two HBM round trips).  ``cpu_baseline`` times the NumPy/SciPy
restatement of the reference path (reference-order float32 propagator,
complex64 pocketfft) on a bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
METRIC = "1024² AS propagations/s & %HBM roofline; served p50/p99 latency"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--batch", type=int, default=64)
    p.add_argument("--n", type=int, default=1024)
    p.add_argument("--phase-mode", default="ref32", choices=["ref32", "accurate"])
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--summarize", nargs=3, metavar=("TRACE_DIR", "FETCH_DIR", "WRITE_DIR"),
                   help="write the rocprofv3 summaries of a profiled run (kernel-trace --stats, "
                        "--pmc FETCH_SIZE, --pmc WRITE_SIZE; csv output) and exit")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hbm_copy_gbps(torch, L, nbytes=1 << 30, reps=10):
    """libcxm's plain complex64 copy (8-byte lanes, like the FFT passes):
    (read + write bytes) / kernel time, timed with the library's HIP events."""
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.complex64, device="cuda")
    dst = torch.empty_like(src)
    s = L.stream_ptr(src.device)
    for _ in range(3):
        L.check(L.lib().cxm_copy_c64(s, src.data_ptr(), dst.data_ptr(), n))
    torch.cuda.synchronize()
    L.timing_enable(True)
    for _ in range(reps):
        L.check(L.lib().cxm_copy_c64(s, src.data_ptr(), dst.data_ptr(), n))
    kt = L.timing_collect()
    L.timing_enable(False)
    ms, cnt = kt["copy_c64"]
    del src, dst
    return 2 * nbytes / (ms / cnt * 1e-3) / 1e9


def cpu_baseline(args):
    """Reference path restated in NumPy/SciPy (oracle, 'port'), host cores."""
    from oracle import propagation as op

    cores = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(0)
    sample_planes = 4
    u = np.exp(1j * rng.uniform(0, 2 * np.pi, (sample_planes, args.n, args.n))).astype(np.complex64)
    op.propagate(u[:1], (0.3, 0.3), 0.532, 100.0, 1.33, 0, precision="ref32", workers=cores)
    done, t0 = 0, time.perf_counter()
    while True:
        op.propagate(u, (0.3, 0.3), 0.532, 100.0, 1.33, 0, precision="ref32", workers=cores)
        done += sample_planes
        el = time.perf_counter() - t0
        if el >= args.cpu_seconds:
            break
    return {
        "value": done / el,
        "unit": "propagations/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{done} planes of {args.n}^2 in calls of {sample_planes} "
                  f"(asm_propagate, reference-order float32 propagator, complex64 "
                  f"scipy.fft workers={cores}), {el:.1f} s",
    }


COPY_BYTES = 1 << 30  # bytes read (and written) by one hbm_copy_gbps launch


def timing_name(rocprof_name: str) -> str | None:
    """rocprof kernel name -> libcxm timing name (col_prop_1024, ...)."""
    import re

    m = re.search(r"col_prop_kernel<(\d+),", rocprof_name)
    if m:
        return f"col_prop_{m.group(1)}"
    m = re.search(r"row_fft_kernel<(\d+), \d+, (-?1)>", rocprof_name)
    if m:
        return f"row_fft_{'fwd' if m.group(2) == '-1' else 'inv'}_{m.group(1)}"
    if "copy_kernel" in rocprof_name:
        return "copy_c64"
    return None


def summarize(trace_dir: str, fetch_dir: str, write_dir: str) -> None:
    """Kernel-time and traffic summaries of a profiled bench run.  Counter
    values (KiB per dispatch) are scaled by what the same counter reads for
    libcxm's copy kernel of known size (COPY_BYTES read, COPY_BYTES written)."""
    import csv
    import glob

    def one(pattern):
        files = glob.glob(os.path.join(pattern))
        if not files:
            raise SystemExit(f"no file matches {pattern}")
        return files[0]

    stats = list(csv.DictReader(open(one(os.path.join(trace_dir, "*kernel_stats.csv")))))
    lines = ["| kernel | calls | avg us | min us | max us | % |", "|---|---|---|---|---|---|"]
    for r in stats:
        lines.append(f"| `{r['Name'][:80]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    counters = {}
    for tag, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        for r in csv.DictReader(open(one(os.path.join(d, "*counter_collection.csv")))):
            name = timing_name(r["Kernel_Name"])
            if name is None or r["Counter_Name"] != tag:
                continue
            counters.setdefault(name, {}).setdefault(tag, []).append(float(r["Counter_Value"]) * 1024)
    avg = {k: {t: sum(v) / len(v) for t, v in d.items()} for k, d in counters.items()}
    cal = avg.get("copy_c64", {})
    f_fetch = COPY_BYTES / cal["FETCH_SIZE"] if cal.get("FETCH_SIZE") else None
    f_write = COPY_BYTES / cal["WRITE_SIZE"] if cal.get("WRITE_SIZE") else None
    traffic = {}
    tl = ["| kernel | FETCH_SIZE raw MiB | WRITE_SIZE raw MiB | scaled read MiB | scaled write MiB |",
          "|---|---|---|---|---|"]
    for k, d in sorted(avg.items()):
        fr, wr = d.get("FETCH_SIZE"), d.get("WRITE_SIZE")
        if fr is None or wr is None or f_fetch is None or f_write is None:
            continue
        rd, wt = fr * f_fetch, wr * f_write
        traffic[k] = rd + wt
        tl.append(f"| {k} | {fr / 2**20:.1f} | {wr / 2**20:.1f} | {rd / 2**20:.1f} | {wt / 2**20:.1f} |")
    print("## Kernel time (rocprofv3 --kernel-trace --stats)\n")
    print("\n".join(lines))
    print("\n## HBM traffic per dispatch (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)\n")
    print(f"Scale factors from the {COPY_BYTES >> 20} MiB copy kernel: read x{f_fetch}, write x{f_write}\n")
    print("\n".join(tl))
    (ROOT / "profiles").mkdir(exist_ok=True)
    (ROOT / "profiles" / "traffic_m0.json").write_text(json.dumps(
        {"bytes_per_launch": traffic, "scale_fetch": f_fetch, "scale_write": f_write}, indent=1) + "\n")


def main():
    args = parse()
    if args.summarize:
        summarize(*args.summarize)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1"}))
        log("bench.py: multi-GPU runs are not part of this build (one GPU per box); use --gpus 1")
        sys.exit(2)

    import torch

    import chromatix_amd as cx
    import chromatix_amd.functional as cf
    from chromatix_amd import _lib as L

    B, N = args.batch, args.n
    log(f"[bench] building input {B}x{N}x{N} complex64")
    rng = np.random.default_rng(0)
    phi = torch.empty((B, N, N), dtype=torch.float32, device="cuda")
    for b in range(B):  # plane by plane to bound host memory
        phi[b] = torch.from_numpy(rng.uniform(0, 2 * np.pi, (N, N)).astype(np.float32))
    u = torch.polar(torch.ones_like(phi), phi)
    del phi
    field = cx.ScalarField(u, 0.3, 0.0, 0.532)

    def step():
        return cf.asm_propagate(field, 100.0, 1.33, 0, mode="same", phase_mode=args.phase_mode)

    for _ in range(args.warmup):
        out = step()
    torch.cuda.synchronize()
    log("[bench] warmup done")

    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(args.steps):
        out = step()
    stop.record()
    torch.cuda.synchronize()
    ms_total = start.elapsed_time(stop)
    ms_step = ms_total / args.steps
    props_s = B * args.steps / (ms_total * 1e-3)
    log(f"[bench] timed {args.steps} steps: {ms_step:.3f} ms/step, {props_s:,.0f} props/s")

    # live per-kernel timing (HIP events around every launch on its stream)
    L.timing_enable(True)
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    kt = L.timing_collect()
    L.timing_enable(False)
    del out
    for k, (ms, cnt) in sorted(kt.items(), key=lambda kv: -kv[1][0]):
        log(f"[bench] kernel {k}: {cnt} launches, avg {ms / cnt * 1e3:.1f} us")
    dom, (dom_ms, dom_cnt) = max(kt.items(), key=lambda kv: kv[1][0])
    dom_avg_s = dom_ms / dom_cnt * 1e-3
    planes_per_launch = B * args.steps / dom_cnt
    # each pass reads and writes every complex64 value of its planes once
    dom_bytes = planes_per_launch * 16 * N * N
    achieved = dom_bytes / dom_avg_s / 1e9
    gpu_ms_per_step = sum(ms for ms, _ in kt.values()) / args.steps
    path_bytes = B * 32 * N * N
    path_achieved = path_bytes / (ms_step * 1e-3) / 1e9

    traffic = None
    tf = ROOT / "profiles" / "traffic_m0.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text())["bytes_per_launch"].get(dom)
        except Exception:
            traffic = None

    copy_gbps = hbm_copy_gbps(torch, L)
    log(f"[bench] device copy: {copy_gbps:,.0f} GB/s")

    result = {
        "metric": METRIC,
        "value": props_s,
        "unit": "propagations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic",
        "config": {
            "workload": f"asm_propagate batch of {B} x {N}x{N} complex64 exp(i*phi), phi~U[0,2pi) "
                        "seed 0, dx 0.3 um, lambda 0.532 um, n 1.33, z 100 um, pad 0, mode same",
            "batch": B,
            "grid": N,
            "phase_mode": args.phase_mode,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dom_bytes,
            "avg_launch_us": dom_avg_s * 1e6,
        },
        "path_roofline": {
            "model": "32*N^2 bytes per propagation (two HBM round trips)",
            "achieved": path_achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": path_achieved / HBM_PEAK_GBPS,
            "gpu_kernel_ms_per_step": gpu_ms_per_step,
            "kernels": {k: {"launches": c, "avg_us": ms / c * 1e3} for k, (ms, c) in kt.items()},
        },
        "hbm_copy_GBps": copy_gbps,
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle port) ...")
        result["cpu_baseline"] = cpu_baseline(args)
        log(f"[bench] CPU: {result['cpu_baseline']['value']:.1f} props/s on "
            f"{result['cpu_baseline']['cores']} cores")
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
