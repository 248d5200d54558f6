"""Benchmark: closed-form alpha solve at 1920x1080 on one MI355X.

One step = one complete alpha solve of the 1920x1080 workload with inputs
already in device memory: closed-form operator set-up, right-hand side,
V-cycle hierarchy (Galerkin coarse operators), PCG to rtol 1e-7 and the
final clip. `value` is the mixed-precision (float32 stencil/SpMV/smoothing,
float64 dots and refinement) throughput in Mpix/s; the float64 PCG is timed
beside it and the two alphas are compared.

The roofline object covers the kernel class that takes most time in the
timed mixed-precision steps, from HIP events recorded around each of its
launches on its own stream: achieved = algorithmic bytes per launch / mean
launch duration. The CPU baseline times the oracle's restatement of the
reference (scipy CSR, vcycle PCG, float64) single-threaded on the 320x240
workload.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per grid point and launch, by kernel class, for vector
# element size s (8 float64, 4 float32): image 3s, class byte 1, coarse
# vectors s/4 per fine point, stored coarse stencils 25s per point.
def bytes_per_point(kind, s):
    return {
        "cf_apply": s + 3 * s + 1 + s,                     # x, I, cls -> A x
        "cf_restrict": s + s + 3 * s + 1 + s / 4,          # b, D, I, cls -> coarse b
        "cf_smooth": s + s + 3 * s + 1 + s / 4 + s,        # b, D, I, cls, e -> z
        "cf_residual": s + s + 3 * s + 1 + s,              # x, b, I, cls -> r
        "dia_apply": 25 * s + 2 * s,
        "dia_restrict": 25 * s + 2 * s + s / 4,
        "dia_smooth": 25 * s + 2 * s + s / 4 + s,
        "dia_residual": 25 * s + 3 * s,
        "cg_update": 6 * s,
        "cg_pupdate": 3 * s,
    }[kind]


KINDS = ["cf_apply", "cf_restrict", "cf_smooth", "cf_residual", "dia_apply", "dia_restrict",
         "dia_smooth", "dia_residual", "cg_update", "cg_pupdate"]


def cpu_baseline():
    """Oracle (port of the reference) on the 320x240 workload, one thread."""
    code = (
        "import sys, time, json; sys.path.insert(0, %r)\n"
        "from oracle.cf import solve_alpha_lambda\n"
        "from pymatting_amd import synthetic\n"
        "image, trimap, _ = synthetic.config1()\n"
        "ts = []\n"
        "for _ in range(3):\n"
        "    t = time.perf_counter(); solve_alpha_lambda(image, trimap); ts.append(time.perf_counter() - t)\n"
        "print(json.dumps({'t': min(ts), 'n': image.shape[0] * image.shape[1]}))\n" % ROOT)
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               HIP_VISIBLE_DEVICES="", CUDA_VISIBLE_DEVICES="")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                         timeout=600)
    if out.returncode != 0:
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": round(r["n"] / r["t"] / 1e6, 4), "unit": "Mpix/s", "cores": 1,
            "kind": "port",
            "sample": "320x240 lambda-system solve (cf_laplacian, make_linear_system, vcycle-PCG "
                      "rtol 1e-7, float64; scipy CSR), best of 3, %.2f s each" % r["t"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the workload is a single-image solve with no "
              "multi-GPU path (see the sharding design notes); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import pymatting_amd as pm
    from pymatting_amd import _ffi, synthetic

    image_h, trimap_h, _ = synthetic.config2()
    h, w = trimap_h.shape
    n = h * w
    image = torch.from_numpy(image_h).cuda()
    trimap = torch.from_numpy(trimap_h).cuda()
    torch.cuda.synchronize()
    L = _ffi.lib()

    def step(precision):
        solver = pm.LambdaSolver(image, trimap, precision=precision)
        alpha = solver.solve(rtol=1e-7)
        return alpha, solver.stats

    for _ in range(args.warmup):
        step("mixed")
        step("f64")
    torch.cuda.synchronize()

    # Timed mixed-precision steps, with per-kernel-class events.
    L.hm_prof_set(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        alpha_m, st_m = step("mixed")
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    prof = {}
    for k, kind in enumerate(KINDS):
        for prec, s in ((0, 8), (1, 4)):
            ms, cnt, pts = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
            L.hm_prof_query(k, prec, ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(pts))
            if cnt.value:
                prof[(kind, prec)] = (ms.value, cnt.value, pts.value, s)
    L.hm_prof_set(0)

    t0 = time.perf_counter()
    for _ in range(args.steps):
        alpha_d, st_d = step("f64")
    torch.cuda.synchronize()
    t_f64 = (time.perf_counter() - t0) / args.steps

    diff = (alpha_m - alpha_d).abs()
    max_d = float(diff.max())
    rms_d = float(diff.pow(2).mean().sqrt())

    (kind, prec), (ms, cnt, pts, s) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_s = ms / cnt / 1e3
    bytes_launch = bytes_per_point(kind, s) * pts / cnt
    achieved = bytes_launch / avg_s / 1e9
    breakdown = {"%s_%s" % (k, "f32" if p else "f64"): {"ms_total": round(v[0], 3),
                                                         "launches": v[1]}
                 for (k, p), v in sorted(prof.items(), key=lambda kv: -kv[1][0])}

    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "alpha Mpix/s at 1080p: FP32+refine vs FP64 PCG, max|Δα|≤1e-4",
        "value": round(n / t_mixed / 1e6, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "1920x1080 lambda-system alpha solve (eps 1e-7, lambda 100, "
                               "V(1,1) vcycle PCG, rtol 1e-7), mixed float32/float64 refinement",
                   "height": h, "width": w, "unknown_fraction": round(float((trimap_h == 0.5).mean()), 4)},
        "fp64": {"value": round(n / t_f64 / 1e6, 3), "unit": "Mpix/s",
                 "ms_per_step": round(t_f64 * 1e3, 3), "iterations": st_d["iterations"]},
        "mixed_iterations": st_m["iterations"],
        "max_abs_dalpha_mixed_vs_fp64": max_d,
        "rms_dalpha_mixed_vs_fp64": rms_d,
        "roofline": {"bound": "hbm", "kernel": "%s_%s" % (kind, "f32" if prec else "f64"),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": int(bytes_launch),
                     "avg_launch_us": round(avg_s * 1e6, 2)},
        "kernel_breakdown_timed_mixed": breakdown,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
