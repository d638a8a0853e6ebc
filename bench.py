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
    p.add_argument("--mbs-concurrency", type=int, default=2,
                   help="c4/c10: volumes solved at once, each on its own stream")
    p.add_argument("--config", default="m0",
                   choices=["m0", "c1", "c2", "c3", "c4", "c5", "c6", "c7", "c8", "c9", "c10"],
                   help="m0: the headline batched 1024^2 AS propagation (default); c1-c10: the "
                        "other workloads (plain AS with padding, 4f PSF + gradient, multislice + "
                        "adjoint, Born solves, serving, scaled-convolution propagation, high-NA "
                        "lens, fluorescent multislice, Microscope imaging, Born solves by "
                        "BiCGStab), one JSON line each")
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


def _cpu_model() -> str:
    """The host CPU's model name with its family / model ids (a virtual
    machine may report only the brand in the name)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (s.strip() for s in line.split(":", 1))
                info.setdefault(k, v)
            elif info:
                break
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    if "cpu family" in info and "model" in info:
        name += f" (family {info['cpu family']}, model {info['model']})"
    return name


def cpu_baseline(args):
    """Reference path restated in NumPy/SciPy (oracle, 'port'), on all host
    cores and (a shorter sample) on one worker."""
    from oracle import propagation as op

    cores = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(0)
    sample_planes = 4
    u = np.exp(1j * rng.uniform(0, 2 * np.pi, (sample_planes, args.n, args.n))).astype(np.complex64)

    def rate(workers, seconds, per_call):
        op.propagate(u[:1], (0.3, 0.3), 0.532, 100.0, 1.33, 0, precision="ref32", workers=workers)
        done, t0 = 0, time.perf_counter()
        while True:
            op.propagate(u[:per_call], (0.3, 0.3), 0.532, 100.0, 1.33, 0, precision="ref32",
                         workers=workers)
            done += per_call
            el = time.perf_counter() - t0
            if el >= seconds:
                return done, el

    done, el = rate(cores, args.cpu_seconds, sample_planes)
    done1, el1 = rate(1, max(2.0, args.cpu_seconds / 3), 1)
    return {
        "value": done / el,
        "unit": "propagations/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "value_1_worker": done1 / el1,
        "sample": f"{done} planes of {args.n}^2 in calls of {sample_planes} "
                  f"(asm_propagate, reference-order float32 propagator, complex64 "
                  f"scipy.fft workers={cores}), {el:.1f} s; 1 worker: {done1} planes, "
                  f"{el1:.1f} s",
    }


COPY_BYTES = 1 << 30  # bytes read (and written) by one hbm_copy_gbps launch


def timing_name(rocprof_name: str) -> str | None:
    """rocprof kernel name -> libcxm timing name (col_prop_1024, ...)."""
    import re

    m = re.search(r"col_prop_(?:persist_)?kernel<(\d+),", rocprof_name)
    if m:
        return f"col_prop_{m.group(1)}"
    m = re.search(r"row_fft_(?:loop_)?kernel<(\d+), \d+, (-?1)>", rocprof_name)
    if m:
        return f"row_fft_{'fwd' if m.group(2) == '-1' else 'inv'}_{m.group(1)}"
    if "h_table_kernel" in rocprof_name:
        return "h_table"
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



# ---------------------------------------------------------------------------
# Other workloads (--config c1..c8).  Each prints one JSON line in the same
# format; value and unit are the workload's own throughput.
# ---------------------------------------------------------------------------
def _prewarm(torch, fn, seconds=0.5):
    """Untimed repetitions of the step for at least ``seconds`` of GPU work
    before the warmup steps proper: a timed region of tens of milliseconds
    otherwise starts while the GPU clocks are still ramping up (a 50-step
    4f-PSF run measured 1.74 ms per step cold, 0.93 warm)."""
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        fn()
        n += 1
        if n % 8 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()


def _timed(torch, fn, steps, warmup):
    _prewarm(torch, fn)
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def _kernels(torch, L, fn, steps):
    L.timing_enable(True)
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    kt = L.timing_collect()
    L.timing_enable(False)
    return {k: {"launches": c, "avg_us": ms / c * 1e3} for k, (ms, c) in kt.items()}


def _result(args, metric, value, unit, ms_step, steps, warmup, dtype, workload, path_bytes,
            kernels, cpu=None, extra=None):
    achieved = path_bytes / (ms_step * 1e-3) / 1e9
    r = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": workload},
        "path_roofline": {"bytes_per_step": path_bytes, "achieved": achieved,
                          "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                          "kernels": kernels},
        "cpu_baseline": cpu,
    }
    if extra:
        r.update(extra)
    return r


def bench_c1(args):
    """512^2 exp(i phi) fields, pad 256 per side (1024^2 FFT grid), intensity
    cropped to 512^2; a batch of 64 fields per step."""
    import torch

    import chromatix_amd as cx
    import chromatix_amd.functional as cf
    from chromatix_amd import _lib as L

    B, n, pad = 64, 512, 256
    rng = np.random.default_rng(0)
    u = torch.polar(torch.ones(B, n, n, device="cuda"),
                    torch.as_tensor(rng.uniform(0, 2 * np.pi, (B, n, n)).astype(np.float32)).cuda())
    f = cx.ScalarField(u, 0.3, 0.0, 0.532)
    step = lambda: cf.propagate_intensity(f, 100.0, 1.33, pad, method="asm", mode="same")
    steps, warm = max(10, args.steps // 4), args.warmup
    ms = _timed(torch, step, steps, warm)
    kern = _kernels(torch, L, step, 5)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import propagation as op

        cores = len(os.sched_getaffinity(0))
        x = np.exp(1j * rng.uniform(0, 2 * np.pi, (n, n))).astype(np.complex64)
        t0, done = time.perf_counter(), 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 8.0):
            np.abs(op.propagate(x, (0.3, 0.3), 0.532, 100.0, 1.33, pad, precision="ref32",
                                workers=cores)) ** 2
            done += 1
        el = time.perf_counter() - t0
        cpu = {"value": done / el, "unit": "propagations/s", "cores": cores, "kind": "port",
               "sample": f"{done} single-field propagations (oracle, ref32 propagator, complex64 "
                         f"scipy.fft workers={cores}), {el:.1f} s"}
    return _result(args, "512^2 AS propagations/s (pad 256, intensity)", B / (ms * 1e-3),
                   "propagations/s", ms, steps, warm, "complex64",
                   f"propagate_intensity batch {B} x 512^2, pad 256 -> 1024^2, z 100 um, same",
                   B * 32 * 1024 * 1024, kern, cpu)


def bench_c2(args):
    """4f PSF stack (1280^2, 11 planes, z in linspace(-5, 5) um) forward and
    the phase-mask gradient of the MSE against a target stack; one step =
    forward (the intensity stack) + gradient, computed together
    (psf_4f_mse_grad(return_intensity=True))."""
    import torch

    import chromatix_amd.functional as cf
    from chromatix_amd import _lib as L

    N, Z = 1280, 11
    lam, nimm, f, NA, ft, pitch = 0.532, 1.33, 8000.0, 0.6, 200000.0, 6.5
    dx = ft * lam / (nimm * N * pitch)
    z = np.linspace(-5, 5, Z)
    rng1, rng2 = np.random.default_rng(1), np.random.default_rng(2)
    phase = torch.as_tensor(rng1.uniform(-np.pi, np.pi, (N, N)).astype(np.float32)).cuda()
    I0 = cf.psf_4f(phase, dx, lam, z, f, nimm, NA, ft)
    target = (I0 * torch.as_tensor(rng2.uniform(0.5, 1.5, (Z, N, N)).astype(np.float32)).cuda())

    def step():
        # the loss's forward (the intensity stack) and its gradient in one
        # sweep, as jax.value_and_grad computes them
        cf.psf_4f_mse_grad(phase, target, dx, lam, z, f, nimm, NA, ft, return_intensity=True)

    steps, warm = max(10, args.steps // 4), args.warmup
    ms = _timed(torch, step, steps, warm)
    kern = _kernels(torch, L, step, 5)
    ms_f = _timed(torch, lambda: cf.psf_4f(phase, dx, lam, z, f, nimm, NA, ft), steps, 2)
    b_f, b_b = Z * 20 * N * N + 4 * N * N, Z * 36 * N * N + 8 * N * N
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import psf as opsf

        t0 = time.perf_counter()
        ph = phase.cpu().numpy().astype(np.float64)
        tg = target.cpu().numpy().astype(np.float64)
        opsf.psf4f_mse_grad(ph, tg, (dx, dx), lam, z, f, nimm, NA, ft)
        el = time.perf_counter() - t0
        cpu = {"value": 1.0 / el, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"one forward + MSE gradient (FP64 NumPy oracle), {el:.1f} s"}
    return _result(args, "4f PSF stack + phase gradient steps/s (1280^2, 11 planes)",
                   1e3 / ms, "steps/s", ms, steps, warm, "complex64",
                   "Optical4FSystemPSF forward (intensity) + d MSE / d phase, 1280^2, Z = 11",
                   b_f + b_b, kern, cpu, {"forward_ms": ms_f})


def bench_c3(args):
    """Multislice 1024^2, 256 slices of 0.5 um, forward and the Delta n
    gradient of an intensity MSE with checkpoints every 16 slices."""
    import torch

    import chromatix_amd as cx
    import chromatix_amd.functional as cf
    from chromatix_amd import _lib as L
    from chromatix_amd.functional import samples as S

    N, D, k, lam, nm, t = 1024, 256, 16, 0.532, 1.33, 0.5
    rng = np.random.default_rng(3)
    g = torch.as_tensor(rng.standard_normal((D, N, N)).astype(np.float32)).cuda()
    # Gaussian smoothing (sigma 2 px) along y and x by FFT, then std 0.01
    fy = torch.fft.fftfreq(N, device="cuda")
    ker = torch.exp(-2 * (np.pi * 2.0) ** 2 * (fy[:, None] ** 2 + fy[None, :] ** 2))
    dn = torch.fft.ifft2(torch.fft.fft2(g) * ker).real
    dn = (dn / dn.std() * 0.01).contiguous()
    ab = torch.zeros_like(dn)
    f = cx.ScalarField(torch.ones((N, N), dtype=torch.complex64, device="cuda"), 0.3, 0.0, lam)
    out = cf.multislice_thick_sample(f, ab, dn, nm, t, 0)
    target = (out.intensity * 1.05).contiguous()
    d = S._desc(f, D, nm, t, 0, None, (0.0, 0.0), None, False, False, False, None, k)
    ck_holder = {}

    def fwd():
        _, ck_holder["ck"] = S._fwd(f.u, dn, ab, d, want_ckpt=True)

    def bwd():
        cf.multislice_mse_grad(f, ab, dn, target, nm, t, 0, checkpoint_every=k,
                               checkpoints=ck_holder["ck"])

    steps, warm = 3, 1
    ms_f = _timed(torch, fwd, steps, warm)
    ms_b = _timed(torch, bwd, steps, warm)
    kern = _kernels(torch, L, lambda: (fwd(), bwd()), 1)
    b_f = D * 36 * N * N
    b_b = D * 100 * N * N + (D // k) * 8 * N * N
    return _result(args, "multislice forward + adjoint steps/s (1024^2, 256 slices, k = 16)",
                   1e3 / (ms_f + ms_b), "steps/s", ms_f + ms_b, steps, warm, "complex64",
                   "multislice_thick_sample 1024^2 x 256 slices (t 0.5 um, dn smoothed noise std "
                   "0.01) forward + d MSE(I) / d dn, checkpoints every 16",
                   b_f + b_b, kern, None, {"forward_ms": ms_f, "adjoint_ms": ms_b,
                                           "checkpoint_every": k})


def _mbs_volumes(torch, mbs, n, bc, vols, k0, sp):
    """The c4 volumes: 128^3 of n = 1.33 with 20 spheres (radius 4-12
    voxels, n in U[1.36, 1.40], seeded), a 16-voxel absorbing boundary and
    an x-polarised plane current sheet."""
    for v in range(vols):
        rng = np.random.default_rng(4 + v)
        eps = np.full((n, n, n), 1.33**2, np.complex64)
        idx = np.indices((n, n, n), dtype=np.float32)
        for _ in range(20):
            c = rng.uniform(0.2 * n, 0.8 * n, 3)
            r = rng.uniform(4, 12)
            inside = np.sum((idx - c[:, None, None, None].astype(np.float32)) ** 2, axis=0) <= r * r
            eps[inside] = rng.uniform(1.36, 1.40) ** 2
        sample = mbs.add_absorbing_bc(mbs.Sample(torch.as_tensor(eps).cuda(), sp), (0, 1, 2), bc * sp,
                                      n_background=1.33)
        J = torch.zeros(sample.shape + (3,), dtype=torch.complex64, device="cuda")
        J[:, :, bc + 1, 0] = 1.0
        yield sample, mbs.Source(J, k0)


def bench_c4(args, use_bicgstab=False):
    """Modified Born series: 8 volumes of 128^3 (+16-voxel absorbing
    boundary), 20 spheres each, solved to a relative residual of 1e-4
    (|prec_y - prec_forward(x)| <= 1e-4 |prec_y|: the fixed point's update
    norm, BiCGStab's residual)."""
    import torch

    from chromatix_amd.experimental import modified_born_series as mbs

    lam = 0.532
    k0, sp = 2 * np.pi / lam, lam / 4
    n, bc, vols = 128, 16, 8
    kw = dict(tol=1e-4, use_bicgstab=True) if use_bicgstab else dict(rtol=1e-4)
    # untimed warm-up (library initialisation, first launches of every
    # kernel, the side streams): short solves of the first volume
    for sample, src in _mbs_volumes(torch, mbs, n, bc, 1, k0, sp):
        c = max(1, args.mbs_concurrency)
        for _ in range(max(1, args.warmup)):
            mbs.solve_batch([sample] * c, [src] * c, maxiter=16, concurrency=c, **kw)
        torch.cuda.synchronize()
    # the volumes are independent: solve_batch runs them args.mbs_concurrency
    # at a time, each on its own stream (1: one after another)
    pairs = list(_mbs_volumes(torch, mbs, n, bc, vols, k0, sp))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = mbs.solve_batch([p[0] for p in pairs], [p[1] for p in pairs], maxiter=3000,
                          concurrency=args.mbs_concurrency, return_info=True, **kw)
    torch.cuda.synchronize()
    times = [(time.perf_counter() - t0) / vols] * vols
    iters = [r.iterations for r in res]
    N3 = (n + 2 * bc) ** 3
    per_it = sum(times) / sum(iters)
    if use_bicgstab:
        # per iteration: two operator applications (V folded into the vector
        # kernels; 3D FFT = 3 passes of 48 B/voxel each way, Green 48 B) and
        # the five vector kernels (128 + 104 + 104 + 80 + 168 B/voxel)
        per_iter_bytes = 2 * (2 * 144 + 48) + 584
        return _result(args, "modified Born series BiCGStab solves/s (128^3 + boundary, tol 1e-4)",
                       vols / sum(times), "solves/s", sum(times) / vols * 1e3, vols, max(1, args.warmup),
                       "complex64",
                       "8 volumes 160^3 (128^3 + 16-voxel boundary), 20 spheres, plane-wave current "
                       "sheet, BiCGStab to relative residual 1e-4",
                       per_iter_bytes * N3 * (sum(iters) / vols), {}, None,
                       {"iterations": iters, "ms_per_iteration": per_it * 1e3,
                        "concurrent_solves": args.mbs_concurrency})
    return _result(args, "modified Born series solves/s (128^3 + boundary, rtol 1e-4)",
                   vols / sum(times), "solves/s", sum(times) / vols * 1e3, vols, max(1, args.warmup),
                   "complex64",
                   "8 volumes 160^3 (128^3 + 16-voxel boundary), 20 spheres, plane-wave current "
                   "sheet, to relative residual 1e-4",
                   272 * N3 * (sum(iters) / vols), {}, None,
                   {"iterations": iters, "ms_per_iteration": per_it * 1e3,
                    "concurrent_solves": args.mbs_concurrency})


def bench_c10(args):
    """bench_c4's volumes solved with BiCGStab (solve(use_bicgstab=True))."""
    return bench_c4(args, use_bicgstab=True)


def bench_c5(args):
    """Serving: Poisson arrivals of a 70/20/10 mix (propagation / SAS+CZT /
    Born), engine vs the plain per-request path, 1000 requests per arrival
    rate for both; p99 against rate for each path and the throughput each
    sustains within a p99 bound."""
    import torch

    from chromatix_amd import serving as sv

    count = int(os.environ.get("CXM_C5_REQUESTS", "1000"))
    rates = [float(r) for r in os.environ.get("CXM_C5_RATES", "25,50,100,150,200,300").split(",")]
    gen = torch.Generator(device="cuda").manual_seed(0)

    def field(shape):
        return torch.polar(torch.ones(shape, device="cuda"),
                           torch.rand(shape, device="cuda", generator=gen) * 6.2832)

    # input pools: requests of one shape share a few input tensors (the
    # results do not depend on which; 200 SAS inputs would be 25 GiB)
    pool_prop = {N: [(field((N, N)), (torch.rand((N, N), device="cuda", generator=gen) - 0.5) * 6.2832)
                     for _ in range(2)] for N in (256, 512, 1024, 2048)}
    wl = np.linspace(0.45, 0.65, 8)
    pool_sas = [field((2048, 2048, 8)) for _ in range(2)]
    # Born volumes: a 12^3 cube of n = 1.38 in n = 1.33, with the absorbing
    # boundary of add_absorbing_bc (8 voxels for 64^3, 16 otherwise), the
    # C4 geometry at three sizes
    from chromatix_amd.experimental import modified_born_series as mbs

    pool_born = {}
    for n in (64, 96, 128):
        bc = 8 if n == 64 else 16
        inner = torch.full((n,) * 3, 1.33**2, dtype=torch.complex64)
        c = n // 2
        inner[c - 6:c + 6, c - 6:c + 6, c - 6:c + 6] = 1.38**2
        eps = mbs.add_absorbing_bc(mbs.Sample(inner, 0.133), (0, 1, 2), bc * 0.133,
                                   n_background=1.33).permittivity.cuda()
        J = torch.zeros(eps.shape + (3,), dtype=torch.complex64, device="cuda")
        J[:, :, bc + 1, 0] = 1
        pool_born[n] = (eps, J)

    def mix(seed, rate):
        rng = np.random.default_rng(seed)
        t, reqs = 0.0, []
        for i in range(count):
            t += rng.exponential(1.0 / rate)
            kind = rng.choice(3, p=[0.7, 0.2, 0.1])
            if kind == 0:
                N = int(rng.choice([256, 512, 1024, 2048]))
                Z = int(rng.choice([1, 4, 16, 32]))
                u, ph = pool_prop[N][i % 2]
                reqs.append(sv.PropRequest(u, ph, rng.uniform(-20, 20, Z), 0.3, 0.532, 1.33, arrival=t))
            elif kind == 1:
                reqs.append(sv.SasCztRequest(pool_sas[i % 2], 0.5, wl, float(rng.uniform(-20, 20)), 1.33,
                                             512, 2.0, arrival=t))
            else:
                eps, J = pool_born[int(rng.choice([64, 96, 128]))]
                reqs.append(sv.BornRequest(eps, J, 0.133, 2 * np.pi / 0.532, 1e-4, arrival=t,
                                           use_bicgstab=True))
        return reqs

    # service time of each request class alone (plain path, after a warm-up)
    alone = {}
    for name, r in (("prop 2048^2 Z=32", sv.PropRequest(pool_prop[2048][0][0], pool_prop[2048][0][1],
                                                        np.linspace(-20, 20, 32), 0.3, 0.532, 1.33)),
                    ("SAS+CZT 2048^2 x 8", sv.SasCztRequest(pool_sas[0], 0.5, wl, 15.0, 1.33, 512, 2.0)),
                    ("Born 64^3+16", sv.BornRequest(*pool_born[64], 0.133, 2 * np.pi / 0.532, 1e-4,
                                                     use_bicgstab=True)),
                    ("Born 96^3+32", sv.BornRequest(*pool_born[96], 0.133, 2 * np.pi / 0.532, 1e-4,
                                                     use_bicgstab=True)),
                    ("Born 128^3+32", sv.BornRequest(*pool_born[128], 0.133, 2 * np.pi / 0.532, 1e-4,
                                                      use_bicgstab=True))):
        sv.run_plain([r])
        t0 = time.perf_counter()
        sv.run_plain([r])
        alone[name] = {"ms": (time.perf_counter() - t0) * 1e3, "iterations": getattr(r, "iterations", None)}
        log(f"[bench] alone: {name} {alone[name]}")

    eng = sv.ServingEngine(max_batch=8, use_graphs=True)
    eng.warm(mix(5, 1.0))
    torch.cuda.synchronize()
    sweep = {}
    for rate in rates:
        st = sv.latency_stats(eng.run(mix(5, rate)))
        sp = sv.latency_stats(sv.run_plain(mix(5, rate)))
        sweep[str(rate)] = {"serving": st, "plain": sp}
        log(f"[bench] rate {rate}/s ({count} requests): serving {st['throughput_rps']:.1f} rps "
            f"p50 {st['p50_ms']:.1f} p99 {st['p99_ms']:.1f} ms | plain {sp['throughput_rps']:.1f} rps "
            f"p50 {sp['p50_ms']:.1f} p99 {sp['p99_ms']:.1f} ms")
    # the throughput each path sustains with its p99 within a bound: the
    # largest measured throughput among the rates whose p99 meets it
    slo = {}
    for bound in (100.0, 250.0, 500.0, 1000.0, 2000.0):
        best = {}
        for path in ("serving", "plain"):
            ok = [v[path]["throughput_rps"] for v in sweep.values() if v[path]["p99_ms"] <= bound]
            best[path] = max(ok) if ok else None
        ratio = best["serving"] / best["plain"] if best["serving"] and best["plain"] else None
        slo[str(int(bound))] = {"serving_rps": best["serving"], "plain_rps": best["plain"], "ratio": ratio}
        log(f"[bench] p99 <= {bound:.0f} ms: serving {best['serving']} rps, plain {best['plain']} rps")
    top = sweep[str(rates[-1])]
    sat_ratio = top["serving"]["throughput_rps"] / top["plain"]["throughput_rps"]
    return _result(args, "served requests/s (70% PSF/AS, 20% SAS+CZT, 10% Born), p50/p99 latency",
                   top["serving"]["throughput_rps"], "requests/s",
                   top["serving"]["span_s"] * 1e3 / top["serving"]["n"], top["serving"]["n"], 0,
                   "complex64",
                   f"{count} Poisson arrivals per rate (seed 5) at {rates} per s; value: the engine "
                   f"at {rates[-1]}/s", 0, {}, None,
                   {"sweep": sweep, "p99_bound": slo, "saturation_ratio": sat_ratio, "alone_ms": alone})


def bench_c6(args):
    """kernel_propagate's scaled-convolution branch (Yu et al.): 16 planes of
    1024^2, ASM propagator, output spacing 0.45 um (input 0.3 um)."""
    import torch

    import chromatix_amd as cx
    import chromatix_amd.functional as cf
    from chromatix_amd import _lib as L

    B, N = 16, 1024
    rng = np.random.default_rng(0)
    u = torch.polar(torch.ones(B, N, N, device="cuda"),
                    torch.as_tensor(rng.uniform(0, 2 * np.pi, (B, N, N)).astype(np.float32)).cuda())
    f = cx.ScalarField(u, 0.3, 0.0, 0.532)
    H = cf.compute_asm_propagator(f, 100.0, 1.33)
    step = lambda: cf.kernel_propagate(f, H, output_dx=0.45, use_czt=False)
    steps, warm = max(10, args.steps // 10), args.warmup
    ms = _timed(torch, step, steps, warm)
    kern = _kernels(torch, L, step, 3)
    # per plane: shifted fft2 (two passes), propagator multiply, two axis passes
    per_plane = 32 * N * N + 24 * N * N + 32 * N * N
    return _result(args, "scaled-convolution propagations/s (1024^2 -> 1024^2 at 1.5x spacing)",
                   B / (ms * 1e-3), "propagations/s", ms, steps, warm, "complex64",
                   f"kernel_propagate(use_czt=False, output_dx 0.45) batch {B} x {N}^2, ASM z 100 um",
                   B * per_plane, kern)


def bench_c7(args):
    """high_na_ff_lens: 16 planes of 1024^2 -> 512^2 (NA 0.8, f 100 um)."""
    import torch

    import chromatix_amd as cx
    import chromatix_amd.functional as cf
    from chromatix_amd import _lib as L

    B, N, M = 16, 1024, 512
    rng = np.random.default_rng(0)
    u = torch.polar(torch.ones(B, N, N, device="cuda"),
                    torch.as_tensor(rng.uniform(0, 2 * np.pi, (B, N, N)).astype(np.float32)).cuda())
    f = cx.ScalarField(u, 0.3, 0.0, 0.532)
    step = lambda: cf.high_na_ff_lens(f, 100.0, 1.33, 0.8, output_shape=(M, M), output_dx=0.2)
    steps, warm = max(10, args.steps // 10), args.warmup
    ms = _timed(torch, step, steps, warm)
    kern = _kernels(torch, L, step, 3)
    per_plane = 16 * N * N + (8 * N * N + 8 * M * N) + (8 * M * N + 8 * M * M)
    return _result(args, "high-NA lens fields/s (1024^2 -> 512^2)", B / (ms * 1e-3), "fields/s", ms,
                   steps, warm, "complex64",
                   f"high_na_ff_lens batch {B} x {N}^2 -> {M}^2, NA 0.8, f 100 um, output dx 0.2 um",
                   B * per_plane, kern)


def bench_c8(args):
    """fluorescent_multislice_thick_sample: 512^2 sample, pad 64 (640^2
    grid), 64 slices, 4 Monte-Carlo samples per step."""
    import torch

    import chromatix_amd as cx
    import chromatix_amd.functional as cf
    from chromatix_amd import _lib as L

    N, pad, D, S = 512, 64, 64, 4
    Np = N + 2 * pad
    rng = np.random.default_rng(0)
    u = torch.ones(N, N, dtype=torch.complex64, device="cuda")
    fl = torch.as_tensor(rng.random((D, N, N)).astype(np.float32)).cuda()
    dn = torch.as_tensor((0.01 * rng.standard_normal((D, N, N))).astype(np.float32)).cuda()
    f = cx.ScalarField(u, 0.3, 0.0, 0.532)
    step = lambda: cf.fluorescent_multislice_thick_sample(f, fl, dn, 1.33, 0.5, pad, key=1,
                                                          num_samples=S)
    steps, warm = 3, 1
    ms = _timed(torch, step, steps, warm)
    kern = _kernels(torch, L, step, 1)
    # per sample and slice: forward row + column pass (32 Np^2) and the slice
    # arrays (12 N^2); backward multiply (24 Np^2), row and column inverse
    # passes (24 Np^2) and the intensity accumulation (8 N^2)
    per = S * D * (80 * Np * Np + 20 * N * N)
    return _result(args, "fluorescent multislice stacks/s (512^2, 64 slices, 4 samples)",
                   1.0 / (ms * 1e-3), "stacks/s", ms, steps, warm, "complex64",
                   f"fluorescent_multislice_thick_sample {N}^2 pad {pad}, {D} slices, {S} samples",
                   per, kern, None, {"ms_per_sample_slice": ms / (S * D)})


def bench_c9(args):
    """Microscope imaging of a 128-plane volume (the reference's single-GPU
    imaging example, examples/parallel_imaging.py:10-57): Optical4FSystemPSF
    at 1536^2 (spacing 0.3 um, zero phase mask, tube lens f 100 um), 128
    defocus planes z in linspace(-4, 4) um, objective f 100 um, n 1.33,
    NA 0.8, lambda 0.532 um; a volume of ones convolved plane by plane with
    the PSF intensity, summed over the planes by the sensor."""
    import torch

    from chromatix_amd import _lib as L
    from chromatix_amd.elements import BasicSensor
    from chromatix_amd.systems import Microscope, Optical4FSystemPSF

    D, N = 128, 1536
    shape = (N, N)
    mic = Microscope(Optical4FSystemPSF(shape, 0.3, 100.0, torch.zeros(shape, device="cuda")),
                     BasicSensor(shape, 0.3, resampling_method=None, reduce_axis=0),
                     f=100.0, n=1.33, NA=0.8, spectrum=0.532)
    volume = torch.ones((D, N, N), device="cuda")
    z = np.linspace(-4, 4, D)
    step = lambda: mic(volume, z)
    steps, warm = 5, 1
    ms = _timed(torch, step, steps, warm)
    psf = mic._process_psf(mic.psf(z))
    img_step = lambda: mic.image(volume, psf)
    ms_img = _timed(torch, img_step, steps, warm)
    kern = _kernels(torch, L, step, 1)
    # the planewise convolution summed over the planes (H = W = 1536 sample
    # and PSF planes, 3072-point transforms, 1536-row planes of 3072 column
    # slots in the workspace), per plane: the row pass reads the sample and
    # PSF planes (4HW + 4HW) and writes 8HL; the column pass reads 8HL and
    # reads and writes the running sum (8HL + 8HL); once per image the
    # inverse row pass reads the sum (8HL) and writes the image (4HW)
    Lc = 3072
    conv_bytes = D * (8 * N * N + 32 * N * Lc) + 8 * N * Lc + 4 * N * N
    return _result(args, "Microscope images/s (128 x 1536^2 4f PSF + planewise Fourier convolution "
                         "+ plane sum)", 1.0 / (ms * 1e-3), "images/s", ms, steps, warm, "float32",
                   f"Microscope(Optical4FSystemPSF {N}^2, {D} planes) on a {D}x{N}^2 volume, "
                   "BasicSensor reduce_axis 0", conv_bytes, kern, None,
                   {"ms_per_image_conv_and_sensor": ms_img,
                    "conv_roofline_frac": conv_bytes / (ms_img * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                    "reference_published_ms": {"value": 172.86, "hardware": "1x NVIDIA A100",
                                               "source": "docs/parallelism.md:182"}})


CONFIGS = {"c1": bench_c1, "c2": bench_c2, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5,
           "c6": bench_c6, "c7": bench_c7, "c8": bench_c8, "c9": bench_c9, "c10": bench_c10}


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

    if args.config != "m0":
        print(json.dumps(CONFIGS[args.config](args)), flush=True)
        return

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

    _prewarm(torch, step)
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
            "traffic_source": "FETCH_SIZE + WRITE_SIZE per dispatch of this kernel from a "
                              "rocprofv3 --pmc run of this workload (profiles/traffic_m0.json), "
                              "each scaled by the counter's reading for a copy kernel of known "
                              "size; not measured in this run",
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
