"""Benchmark: closed-form alpha solve at 1920x1080 on one MI355X.

One step = one complete alpha solve of the 1920x1080 workload (the
lambda-penalised closed-form system, eps 1e-7, lambda 100, V(1,1) V-cycle
PCG) with the image and trimap already in device memory: operator set-up,
right-hand side, V-cycle hierarchy with Galerkin coarse operators, PCG and
the final clip.

`value` is the throughput of the mixed form: one CG whose recurrences run
in float32 (stencil, SpMV, V-cycle smoothing) with float64 dot products,
the solution accumulated in float64 and the recursive residual replaced by
the float64 residual b - A x after every drop by 0.01 (CG with reliable
updates; convergence confirmed on a float64 residual). It runs at the
reference's default rtol 1e-7, and its alpha is compared with the float64
PCG's at the same rtol 1e-7 (the metric's max|d alpha| <= 1e-4 bound; the
line says whether it holds). At rtol 1e-7 both forms stop about 3e-3 from
the converged alpha (float64, rtol 1e-11, computed outside the timed
regions): that is the reference's own CG truncation, reported beside the
value together with a strict pair (both forms at rtol 1e-9, each within
1e-4 of the converged alpha) and the restarted float32-in-float64
refinement.

The roofline object covers the longest kernel of the value form's
iteration, the finest grid's pre-smooth + residual + full-weighting
restriction (13 float32 coefficient planes, r and 1/D in, the coarse
right-hand side out): HIP events on the library's stream bracket 50
back-to-back launches on the solve's own V-cycle, and achieved = algorithmic
bytes per launch / mean launch duration. The fine prolongation +
post-smooth and the stored-stencil SpMV y = A x are measured the same way
beside it. Per-class breakdowns of one extra step of each form (events
around every launch, which adds a few microseconds to each) give the
per-grid-level figures and the step's algorithmic bytes. The CPU baseline times the oracle's restatement of the reference
(scipy CSR, vcycle PCG at rtol 1e-7, float64) single-threaded on the same
1920x1080 workload, one solve.

The line also carries the other configurations' figures, measured in the
same run: "c3" (64 images batched vs one at a time), "c4" (3840x2160 with a
5-level V(2,2) cycle, and the float32 foreground against float64) and "c5"
(gradient of a 16x512x512 batch against finite differences and the float64
adjoint), and "other_rows_320x240" (ichol, the foreground methods, large-kernel
matting, radius 2 and the cutout pipeline on the 320x240 image, each beside
its single-threaded CPU oracle). --c2-only leaves them out.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--c2-only]
       python bench.py --pmc-probe   (SpMV launches, V-cycle applications and a
                                      256 MiB copy only, for hardware-counter runs)
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
# HBM bytes per launch of float32 kernels, from two rocprofv3 --pmc passes
# over `python bench.py --pmc-probe` (FETCH_SIZE, then WRITE_SIZE, KiB per
# dispatch), each scaled by what the same counter reads for a 256 MiB copy
# (FETCH_SIZE x 1.999788, WRITE_SIZE x 1.0):
#   1080p fine-grid restriction: 63541.5 KiB x 1.999788 + 2025.0 KiB
#     = 132.2 MB (algorithmic 126.5 MB);
#   1080p fine-grid prolongation + smoothing: 63341.6 KiB x 1.999788
#     + 8100.0 KiB = 138.0 MB (algorithmic 134.8 MB);
#   1080p stored-stencil SpMV: 57550.0 KiB x 1.999788 + 8100.0 KiB
#     = 126.1 MB (algorithmic 124.4 MB);
#   4K stored-stencil SpMV: 233182.9 KiB x 1.999788 + 32400.0 KiB
#     = 510.7 MB (algorithmic 497.7 MB).
TRAFFIC_RESTRICT_F32 = 132192787
TRAFFIC_SMOOTH_F32 = 138004242
TRAFFIC_SPMV_F32 = 126144221
TRAFFIC_SPMV_F32_4K = 510685591

# Algorithmic bytes per grid point and launch, by kernel class, for vector
# element size s (8 float64, 4 float32). Matrix-free passes read the image
# (3s), the class byte and the window mean and inverse covariance (9s per
# window, one window per point); stored stencils read their 13 symmetric
# coefficients (13s); coarse vectors add s/4 per fine point.
def bytes_per_point(kind, s):
    return {
        "cf_apply": 2 * s + s + 3 * s + 9 * s + 1 + 2 * s,  # z, p_old, I, W, cls -> q, p
        "cf_restrict": s + s + 3 * s + 9 * s + 1 + s / 4,   # b, D, I, W, cls -> coarse b
        "cf_smooth": s + s + 3 * s + 9 * s + 1 + s / 4 + s,  # b, D, I, W, cls, e -> z
        "cf_residual": s + s + 3 * s + 9 * s + 1 + s,       # x, b, I, W, cls -> r
        "dia_apply": 13 * s + 2 * s + 2 * s,                # coef, z, p_old -> q, p
        "dia_restrict": 13 * s + 2 * s + s / 4,             # coef, b, D -> coarse b
        "dia_smooth": 13 * s + 2 * s + s / 4 + s,           # coef, b, D, e -> z
        "dia_residual": 13 * s + 3 * s,
        "cg_update": 6 * s,
        "cg_pupdate": 3 * s,
        "tail": None,
        "fg_sweep": None,
        "coarse_apply": 13 * s + 2 * s,
        "coarse_restrict": 13 * s + 2 * s + s / 4,
        "coarse_smooth": 13 * s + 2 * s + s / 4 + s,
        "coarse_residual": 13 * s + 3 * s,
        "cg_update_x64": 8 + s + 2 * s + s + 8,            # x (f64), p, r, q -> x, r
        "cg_update_r": 3 * s,                               # r, q -> r (x beside the tail)
        "l1_apply": 13 * s + 2 * s,
        "l1_restrict": 13 * s + 2 * s + s / 4,
        "l1_smooth": 13 * s + 2 * s + s / 4 + s,
        "l1_residual": 13 * s + 3 * s,
        # Matrix-free passes that recompute the windows from the staged image
        # read no window planes.
        "cfrec_apply": 2 * s + s + 3 * s + 1 + 2 * s,
        "cfrec_restrict": s + s + 3 * s + 1 + s / 4,
        "cfrec_smooth": s + s + 3 * s + 1 + s / 4 + s,
        "cfrec_residual": s + s + 3 * s + 1 + s,
        # Several coarse levels in one launch, per point of the launch's
        # finest level a (a level-a+j point stands for 4^-j of them): the
        # descent pair reads level a's stencil, b and 1/D and level a+1's
        # stencil and 1/D and writes b of levels a+1 and a+2; the ascent chain
        # of four levels reads each level's stencil, b and 1/D and the
        # correction of level a+4 and writes z of level a.
        "pair_down": 13 * s + 2 * s + (13 * s + s + s) / 4 + s / 16,
        "pair_up": (13 * s + 2 * s) * (1 + 1 / 4 + 1 / 16 + 1 / 64) + s + s / 256,
    }[kind]


KINDS = ["cf_apply", "cf_restrict", "cf_smooth", "cf_residual", "dia_apply", "dia_restrict",
         "dia_smooth", "dia_residual", "cg_update", "cg_pupdate", "tail", "fg_sweep",
         "coarse_apply", "coarse_restrict", "coarse_smooth", "coarse_residual", "cg_update_x64",
         "cg_update_r", "l1_apply", "l1_restrict", "l1_smooth", "l1_residual", "cfrec_apply",
         "cfrec_restrict", "cfrec_smooth", "cfrec_residual", "pair_down", "pair_up"]


# Algorithmic bytes per grid point of the stored-stencil SpMV y = A x:
# 13 coefficient planes, x and y.
def spmv_bytes_per_point(s):
    return 13 * s + 2 * s


def profile_classes(L, fn):
    """Runs fn() once with per-launch events and returns {(kind, prec): (ms,
    launches, points, element size)} for every kernel class it launched."""
    import torch
    L.hm_prof_set(1)
    torch.cuda.synchronize()
    fn()
    torch.cuda.synchronize()
    out = {}
    for k, kind in enumerate(KINDS):
        for prec, s in ((0, 8), (1, 4)):
            ms, cnt, pts = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
            L.hm_prof_query(k, prec, ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(pts))
            if cnt.value:
                out[(kind, prec)] = (ms.value, cnt.value, pts.value, s)
    L.hm_prof_set(0)
    return out


def fine_passes(prof, prefix):
    """Per fine-grid pass class (apply, restrict, smooth) of a float32
    level-0 form: algorithmic bytes per point, event-timed mean launch and the
    GB/s they give (events around each launch add a few microseconds)."""
    out = {}
    for cls in ("apply", "restrict", "smooth"):
        v = prof.get(("%s_%s" % (prefix, cls), 1))
        if v is None:
            continue
        ms, cnt, pts, s = v
        bpp = bytes_per_point("%s_%s" % (prefix, cls), s)
        out[cls] = {"bytes_per_point": bpp, "launches": cnt,
                    "avg_us_with_events": round(ms / cnt * 1e3, 2),
                    "GBps_with_events": round(bpp * pts / (ms * 1e-3) / 1e9, 1)}
    return out


def time_spmv(L, handle, prec, x, y, reps=50):
    """Mean duration of `reps` back-to-back hm_dia_apply launches (HIP events
    on the library's stream, which is torch's current stream)."""
    import torch
    from pymatting_amd import _ffi
    stream = _ffi.stream()
    for _ in range(3):
        _ffi.check(L.hm_dia_apply(handle, prec, _ffi.ptr(x), _ffi.ptr(y), stream))
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        _ffi.check(L.hm_dia_apply(handle, prec, _ffi.ptr(x), _ffi.ptr(y), stream))
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3  # seconds


def pmc_probe():
    """20 float32 SpMV launches and 20 V-cycle applications on the 1080p stored
    operator, 20 SpMV launches on the 3840x2160 one, and 5 copies of 256 MiB
    (the calibration kernel for FETCH_SIZE/WRITE_SIZE), nothing else."""
    import torch
    import pymatting_amd as pm
    from pymatting_amd import _ffi, synthetic
    image_h, trimap_h, _ = synthetic.config2()
    solver = pm.LambdaSolver(torch.from_numpy(image_h).cuda(), torch.from_numpy(trimap_h).cuda(),
                             precision="mixed", level0="stencil")
    n = solver.op.n
    x = torch.rand(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    L = _ffi.lib()
    for _ in range(20):
        _ffi.check(L.hm_dia_apply(solver.op.handle, _ffi.HM_F32, _ffi.ptr(x), _ffi.ptr(y),
                                  _ffi.stream()))
    # V-cycle applications: each runs the fine-grid restriction (pre-smooth +
    # residual + full weighting) and prolongation + post-smooth once.
    for _ in range(20):
        solver.M.apply(x, y)
    del solver, x, y
    torch.cuda.empty_cache()
    # The 3840x2160 SpMV, whose 497.7 MB working set is larger than the
    # 256 MB Infinity Cache.
    image_h, trimap_h, _ = synthetic.config4()
    solver = pm.LambdaSolver(torch.from_numpy(image_h).cuda(), torch.from_numpy(trimap_h).cuda(),
                             precision="mixed", level0="stencil")
    n4 = solver.op.n
    x = torch.rand(n4, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(20):
        _ffi.check(L.hm_dia_apply(solver.op.handle, _ffi.HM_F32, _ffi.ptr(x), _ffi.ptr(y),
                                  _ffi.stream()))
    a = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "spmv_f32_1080p", "n": n, "spmv_launches": 20,
                      "vcycle_applications": 20, "n_4k": n4, "spmv_4k_launches": 20,
                      "copy_bytes": a.numel() * 4, "copy_launches": 5}))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(alpha_path=None):
    """The oracle (a CPU restatement of the reference: numpy/scipy CSR assembly,
    vcycle-preconditioned CG at the reference's default rtol 1e-7, float64) on
    the same 1920x1080 workload, once, in one thread. Beside it, the
    reference's own default path on the same image: estimate_alpha_cf (the
    unknown-pixel system with the ichol preconditioner, rtol 1e-7). With
    alpha_path, the oracle's lambda-system alpha is saved there (.npy) for the
    parity figures of the bench line."""
    code = (
        "import sys, time, json; sys.path.insert(0, %r)\n"
        "from oracle.cf import solve_alpha_lambda, estimate_alpha_cf\n"
        "from pymatting_amd import synthetic\n"
        "image, trimap, _ = synthetic.config2()\n"
        "t = time.perf_counter(); a, _ = solve_alpha_lambda(image, trimap); t = time.perf_counter() - t\n"
        "if %r: import numpy as np; np.save(%r, a)\n"
        "t1 = time.perf_counter(); estimate_alpha_cf(image, trimap)\n"
        "t1 = time.perf_counter() - t1\n"
        "print(json.dumps({'t': t, 't_a1': t1, 'n': image.shape[0] * image.shape[1]}))\n"
        % (ROOT, alpha_path, alpha_path))
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               HIP_VISIBLE_DEVICES="", CUDA_VISIBLE_DEVICES="")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                         timeout=900)
    if out.returncode != 0:
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": round(r["n"] / r["t"] / 1e6, 4), "unit": "Mpix/s", "cores": 1,
            "kind": "port",
            "sample": "one 1920x1080 lambda-system solve (cf_laplacian, make_linear_system, "
                      "vcycle-PCG rtol 1e-7, float64; numpy/scipy CSR), %.1f s" % r["t"],
            "estimate_alpha_cf_ichol": {"value": round(r["n"] / r["t_a1"] / 1e6, 4),
                                        "unit": "Mpix/s", "seconds": round(r["t_a1"], 2)},
            "affinity_cores": len(os.sched_getaffinity(0)), "host_cpus": os.cpu_count(),
            "cpu_model": _cpu_model()}


def bench_c3(args, forms_run=("f64", "mixed", "f64_rtol_1e-7", "mixed_rtol_1e-7",
                               "mixed_refine")):
    """64 images, heights and widths uniform in 256..1024 (step 16), unknown
    bands 5-40 px: all alpha solves (set-up + PCG + clip) batched, against a
    loop of one-image solves with the same settings. Float64 PCG and the mixed
    form with reliable updates (delta 0.01; the line's value) run to rtol
    1e-9, where every image is within 1e-4 of its converged alpha. At rtol
    1e-7 (the pair beside, the mixed alphas compared with the float64 ones)
    the two forms drift apart by more than 1e-4 on some images. Restarted
    refinement (rtol 1e-8, inner 1e-3) beside. Tile shapes and iterations per
    graph come from BatchSolver.autotune (run once per form, outside the
    timed steps)."""
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic

    data = synthetic.config3()
    images = [torch.from_numpy(d[0]).cuda() for d in data]
    trimaps = [torch.from_numpy(d[1]).cuda() for d in data]
    npix = sum(int(t.numel()) for t in trimaps)
    forms = {"f64": dict(precision="f64", rtol=1e-9),
             "mixed": dict(precision="mixed", rtol=1e-9, mixed="reliable", delta=0.01),
             "f64_rtol_1e-7": dict(precision="f64", rtol=1e-7),
             "mixed_rtol_1e-7": dict(precision="mixed", rtol=1e-7, mixed="reliable", delta=0.01),
             "mixed_refine": dict(precision="mixed", rtol=1e-8, inner_rtol=1e-3)}

    def batched(form, tuned=None):
        cfg = dict(forms[form])
        bs = pm.BatchSolver(images, trimaps, precision=cfg.pop("precision"))
        if tuned is not None:
            bs.level0_shape, bs.coarse_shape, bs.graph_iters = tuned
        alphas = bs.solve(**cfg)
        return alphas, bs

    def sequential(form):
        cfg = dict(forms[form])
        prec = cfg.pop("precision")
        out, its = [], []
        for im, tr in zip(images, trimaps):
            s = pm.LambdaSolver(im, tr, precision=prec, level0="stencil")
            out.append(s.solve(**cfg))
            its.append(s.stats["iterations"])
        return out, its

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = fn()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) / args.steps

    ref = [pm.LambdaSolver(im, tr, precision="f64", level0="stencil").solve(rtol=1e-11,
                                                                        maxiter=100000)
           for im, tr in zip(images, trimaps)]
    res, batched_alphas = {}, {}
    for form in forms_run:
        tuner = pm.BatchSolver(images, trimaps, precision=forms[form]["precision"])
        cfg = {k: v for k, v in forms[form].items() if k != "precision"}
        choice = tuner.autotune(**cfg)
        tuning = getattr(tuner, "tuning", None)
        del tuner
        (alphas, bs), tb = timed(lambda: batched(form, choice))
        (alphas_s, its_s), ts = timed(lambda: sequential(form))
        err = max(float((a - r).abs().max()) for a, r in zip(alphas, ref))
        err_s = max(float((a - r).abs().max()) for a, r in zip(alphas_s, ref))
        its = [st["iterations"] for st in bs.stats]
        batched_alphas[form] = alphas
        res[form] = {"batched_Mpix_per_s": round(npix / tb / 1e6, 3),
                     "batched_ms": round(tb * 1e3, 3),
                     "sequential_Mpix_per_s": round(npix / ts / 1e6, 3),
                     "sequential_ms": round(ts * 1e3, 3),
                     "speedup": round(ts / tb, 3),
                     "iterations_min_max": [min(its), max(its)],
                     "sequential_iterations_min_max": [min(its_s), max(its_s)],
                     "max_abs_dalpha_vs_converged": err,
                     "sequential_max_abs_dalpha": err_s,
                     "autotuned": {"level0_shape": choice[0], "coarse_shape": choice[1],
                                   "graph_iters": choice[2], "timings_s": tuning},
                     "options": {k: v for k, v in forms[form].items()}}
    # The mixed batches against the float64 batch at the same rtol, image by
    # image.
    for form, f64 in (("mixed", "f64"), ("mixed_rtol_1e-7", "f64_rtol_1e-7")):
        res[form]["max_abs_dalpha_vs_f64"] = max(
            float((a - r).abs().max()) for a, r in zip(batched_alphas[form], batched_alphas[f64]))
    # Launch per step against one persistent kernel: a kernel boundary and a
    # grid-wide barrier (cooperative grid.sync and an atomic counter barrier)
    # for grids of 8, 64 and 256 workgroups.
    from pymatting_amd import _ffi
    L = _ffi.lib()
    sync = {}
    for G in (8, 64, 256):
        lu, cu, au, ok = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _ffi.check(L.hm_time_sync(G, 2000, ctypes.byref(lu), ctypes.byref(cu), ctypes.byref(au),
                                  ctypes.byref(ok), _ffi.stream()))
        sync[str(G)] = {"kernel_boundary_us": round(lu.value, 3),
                        "coop_grid_sync_us": round(cu.value, 3),
                        "counter_barrier_us": round(au.value, 3), "barrier_ok": bool(ok.value)}
    graph_vs_direct = None
    tm = res["mixed"]["autotuned"]["timings_s"]
    if tm and tm.get("graph_iters_s"):
        g = tm["graph_iters_s"]
        graph_vs_direct = {"direct_ms": round(g[0] * 1e3, 3) if 0 in g else None,
                           "graph_ms": round(g[8] * 1e3, 3) if 8 in g else None}
    line = {
        "metric": "batched alpha Mpix/s, 64 images 256..1024 px (mixed form)",
        "value": res["mixed"]["batched_Mpix_per_s"], "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["mixed"]["batched_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "64 lambda-system alpha solves (eps 1e-7, lambda 100, V(1,1), "
                               "rtol 1e-9: every image within 1e-4 of its converged alpha) of "
                               "images with h, w uniform in 256..1024 step 16, bands 5-40 px; "
                               "set-up + batched PCG + clip; mixed form with reliable updates",
                   "images": len(images), "total_pixels": npix},
        "f64": res["f64"], "mixed": res["mixed"], "mixed_refine": res.get("mixed_refine"),
        "rtol_1e-7": {"f64": res["f64_rtol_1e-7"], "mixed": res["mixed_rtol_1e-7"]},
        # Whole batched value-form solves (autotune timings, best of one):
        # each iteration's fixed parts replayed from HIP graphs, or launched
        # directly.
        "value_form_graph_vs_direct": graph_vs_direct,
        "kernel_boundary_vs_grid_barrier": sync,
    }
    return line


def bench_c4(args):
    """3840x2160 (unknown band 40 px): mixed form with a 5-level V-cycle, two
    pre- and two post-smoothing float32 Jacobi sweeps and an approximate
    coarsest solve (240x135 grid, 2 damped Jacobi sweeps). The value: stored
    float32 stencil with reliable updates at rtol 4e-10, within 1e-4 of the
    converged alpha. Beside it: both forms at rtol 1e-7 (there they stop
    ~2e-2 from the converged alpha and several 1e-3 apart), and restarted
    refinement (rtol 1e-7, inner 1e-3) on the stored and on the matrix-free
    stencil. Then estimate_foreground_ml in float32 against float64 on the
    converged alpha."""
    import numpy as np
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic

    image_h, trimap_h, _ = synthetic.config4()
    h, w = trimap_h.shape
    n = h * w
    image = torch.from_numpy(image_h).cuda()
    trimap = torch.from_numpy(trimap_h).cuda()
    base = dict(num_pre_iter=2, num_post_iter=2, max_levels=5, coarse_sweeps=2)
    INNER = 1e-3

    def run(prec, level0, **kw):
        solver = pm.LambdaSolver(image, trimap, precision=prec, level0=level0, **base)
        return solver.solve(**kw), solver.stats

    def timed(prec, level0, **kw):
        for _ in range(args.warmup):
            run(prec, level0, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            alpha, st = run(prec, level0, **kw)
        torch.cuda.synchronize()
        return alpha, st, (time.perf_counter() - t0) / args.steps

    # The value: reliable updates at rtol 4e-10, within 1e-4 of the converged
    # alpha. Beside it: reliable updates and float64 PCG (same V-cycle) at
    # rtol 1e-7, and restarted refinement on the stored and the matrix-free
    # stencils.
    alpha_rs, st_rs, t_rs = timed("mixed", "stencil", rtol=4e-10, mixed="reliable")
    alpha_r, st_r, t_r = timed("mixed", "stencil", rtol=1e-7, mixed="reliable")
    alpha_d, st_d, t_d = timed("f64", "stencil", rtol=1e-7)
    alpha_s, st_s, t_s = timed("mixed", "stencil", rtol=1e-7, inner_rtol=INNER)
    alpha_f, st_f, t_f = timed("mixed", "matrix_free", rtol=1e-7, inner_rtol=INNER)
    # The value form on the two matrix-free level-0 operators.
    alpha_fw, st_fw, t_fw = timed("mixed", "matrix_free", rtol=4e-10, mixed="reliable")
    alpha_fr, st_fr, t_fr = timed("mixed", "matrix_free_recompute", rtol=4e-10, mixed="reliable")
    from pymatting_amd import _ffi
    L = _ffi.lib()
    passes = {}
    for name, level0, prefix in (("stored_stencil", "stencil", "dia"),
                                 ("matrix_free_stored_windows", "matrix_free", "cf"),
                                 ("matrix_free_recomputed_windows", "matrix_free_recompute",
                                  "cfrec")):
        passes[name] = fine_passes(profile_classes(L, lambda: run(
            "mixed", level0, rtol=4e-10, mixed="reliable")), prefix)
    conv = pm.LambdaSolver(image, trimap, precision="f64", level0="stencil")
    alpha_c = conv.solve(rtol=1e-11, maxiter=100000)
    # The float32 stored-stencil SpMV at 4K: 431 MB of coefficients, more
    # than the 256 MB MALL, so every launch streams from HBM.
    xs = torch.rand(n, device="cuda", dtype=torch.float32)
    ys = torch.empty_like(xs)
    t_spmv = time_spmv(_ffi.lib(), conv.op.handle, _ffi.HM_F32, xs, ys)
    spmv_gbps = spmv_bytes_per_point(4) * n / t_spmv / 1e9

    def err(a, ref=None):
        d = (a - (alpha_c if ref is None else ref)).abs()
        return float(d.max()), float(d.pow(2).mean().sqrt())

    # Foreground on the converged alpha: float32 (the reference's typing)
    # against float64.
    a_np = alpha_c.clamp(0, 1)
    img32, a32 = image.float(), a_np.float()
    for _ in range(args.warmup):
        pm.estimate_foreground_ml(img32, a32, return_background=True, precision="f32")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        F32, B32 = pm.estimate_foreground_ml(img32, a32, return_background=True,
                                             precision="f32")
    torch.cuda.synchronize()
    t_fg = (time.perf_counter() - t0) / args.steps
    F64, B64 = pm.estimate_foreground_ml(image, a_np, return_background=True, precision="f64")
    dF = float((F32.double() - F64).abs().max())
    dB = float((B32.double() - B64).abs().max())
    line = {
        "metric": "alpha Mpix/s at 4K: mixed 5-level V(2,2) PCG, stored FP32 stencil",
        "value": round(n / t_rs / 1e6, 3), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_rs * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "3840x2160 lambda-system alpha solve, unknown band 40 px, mixed "
                               "float32/float64, 5-level V-cycle (2 pre, 2 post sweeps, coarsest "
                               "240x135 by 2 Jacobi sweeps), reliable updates (float32 CG, "
                               "float64 solution and residual replacement), rtol 4e-10: within "
                               "1e-4 of the converged alpha",
                   "height": h, "width": w, "rtol": 4e-10},
        "stored_fp32_reliable_updates": {"rtol": 4e-10, "delta": 0.01,
                                         "iterations": st_rs["iterations"],
                                         "max_abs_dalpha_vs_converged": err(alpha_rs)[0],
                                         "rms_dalpha_vs_converged": err(alpha_rs)[1]},
        "rtol_1e-7": {
            "mixed": {"value": round(n / t_r / 1e6, 3), "ms_per_step": round(t_r * 1e3, 3),
                      "iterations": st_r["iterations"],
                      "max_abs_dalpha_vs_f64": err(alpha_r, alpha_d)[0],
                      "max_abs_dalpha_vs_converged": err(alpha_r)[0]},
            "f64": {"value": round(n / t_d / 1e6, 3), "ms_per_step": round(t_d * 1e3, 3),
                    "iterations": st_d["iterations"],
                    "max_abs_dalpha_vs_converged": err(alpha_d)[0]}},
        "stored_fp32_refinement": {"value": round(n / t_s / 1e6, 3),
                                   "ms_per_step": round(t_s * 1e3, 3), "rtol": 1e-7,
                                   "inner_rtol": INNER, "iterations": st_s["iterations"],
                                   "max_abs_dalpha_vs_f64": err(alpha_s, alpha_d)[0],
                                   "max_abs_dalpha_vs_converged": err(alpha_s)[0]},
        "matrix_free_fp32_refinement": {"value": round(n / t_f / 1e6, 3),
                                        "ms_per_step": round(t_f * 1e3, 3), "rtol": 1e-7,
                                        "iterations": st_f["iterations"],
                                        "max_abs_dalpha_vs_converged": err(alpha_f)[0]},
        "level0_forms_reliable_rtol_4e-10": {
            "stored_stencil": {"value": round(n / t_rs / 1e6, 3),
                               "iterations": st_rs["iterations"],
                               "max_abs_dalpha_vs_converged": err(alpha_rs)[0]},
            "matrix_free_stored_windows": {"value": round(n / t_fw / 1e6, 3),
                                           "iterations": st_fw["iterations"],
                                           "max_abs_dalpha_vs_converged": err(alpha_fw)[0]},
            "matrix_free_recomputed_windows": {"value": round(n / t_fr / 1e6, 3),
                                               "iterations": st_fr["iterations"],
                                               "max_abs_dalpha_vs_converged": err(alpha_fr)[0]},
            "fine_passes_f32": passes},
        "foreground_f32": {"ms": round(t_fg * 1e3, 3), "Mpix_per_s": round(n / t_fg / 1e6, 2),
                           "max_abs_dF_vs_f64": dF, "max_abs_dB_vs_f64": dB},
        "spmv_f32_4k": {"bytes_per_launch": int(spmv_bytes_per_point(4) * n),
                        "traffic": TRAFFIC_SPMV_F32_4K,
                        "avg_launch_us": round(t_spmv * 1e6, 2),
                        "achieved_GBps": round(spmv_gbps, 1),
                        "frac": round(spmv_gbps / HBM_PEAK_GBS, 4)},
    }
    return line


def bench_c5(args):
    """Gradient of mean((alpha - target)^2) over 16 images of 512x512 (float32
    tensors) through the batched mixed-precision solve and its hand-written
    adjoint (one batched PCG forward, one batched adjoint PCG backward); the
    same gradient with the images solved one at a time is timed beside it.
    Checks: central finite differences (step 1e-3, float64 solves to rtol
    1e-13) on 200 random pixel channels of the unknown regions, and the
    float64 adjoint gradient (float64 solves to rtol 1e-12) as the reference."""
    import numpy as np
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic

    data = synthetic.config5()
    images = np.stack([d[0] for d in data])
    trimaps = np.stack([d[1] for d in data])
    targets = np.stack([d[2] for d in data])
    B, h, w, _ = images.shape
    img32 = torch.from_numpy(images.astype(np.float32)).cuda()
    tri = torch.from_numpy(trimaps).cuda()
    tgt = torch.from_numpy(targets).cuda()

    def forward_backward(img, **opts):
        x = img.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        alpha = pm.solve_alpha(x, tri, **opts)
        loss = ((alpha.double() - tgt) ** 2).mean()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return float(loss.detach()), x.grad.double(), t1 - t0, t2 - t1

    for _ in range(args.warmup):
        forward_backward(img32)
    tf, tb = 0.0, 0.0
    for _ in range(args.steps):
        loss32, g32, f, b = forward_backward(img32)
        tf += f
        tb += b
    tf /= args.steps
    tb /= args.steps
    # The same gradient with the images solved one at a time.
    _, g_loop, tf_loop, tb_loop = forward_backward(img32, batched=False)
    loop_rel = float((g_loop - g32).norm() / g32.norm())
    img64 = torch.from_numpy(images).cuda()
    f64 = dict(precision="f64", rtol=1e-12)
    loss64, g64, _, _ = forward_backward(img64, **f64)
    rel_l2 = float((g32 - g64).norm() / g64.norm())

    fd_opts = dict(precision="f64", rtol=1e-13)
    rng = np.random.default_rng(5)
    step = 1e-3
    fd_err, gmax = [], float(g64.abs().max())
    n_px = B * h * w
    with torch.no_grad():
        def loss_of(i, img_i):
            a = pm.solve_alpha(img_i, tri[i], **fd_opts)
            return float(((a - tgt[i]) ** 2).sum()) / n_px
        for _ in range(200):
            i = int(rng.integers(0, B))
            unk = np.argwhere(trimaps[i] == 0.5)
            y, x = unk[int(rng.integers(0, len(unk)))]
            c = int(rng.integers(0, 3))
            ip = img64[i].clone()
            im = img64[i].clone()
            ip[y, x, c] += step
            im[y, x, c] -= step
            fd = (loss_of(i, ip) - loss_of(i, im)) / (2 * step)
            fd_err.append(abs(fd - float(g64[i, y, x, c])) / gmax)
    line = {
        "metric": "gradient of a 16x512x512 batch: backward/forward time, adjoint vs FD",
        "value": round(tb / tf, 4), "unit": "backward/forward", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round((tf + tb) * 1e3, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "16 images 512x512 float32, loss mean((alpha - target)^2), "
                               "target = true alpha + N(0, 0.05^2); forward: batched mixed "
                               "float32/float64 PCG over the 16 images, rtol 1e-8; backward: "
                               "batched adjoint PCG + image gradient", "batch": B, "height": h,
                   "width": w},
        "forward_ms": round(tf * 1e3, 3), "backward_ms": round(tb * 1e3, 3),
        "one_image_at_a_time": {"forward_ms": round(tf_loop * 1e3, 3),
                                "backward_ms": round(tb_loop * 1e3, 3),
                                "grad_rel_l2_vs_batched": loop_rel},
        "grad_vs_fp64_adjoint_rel_l2": rel_l2,
        "fd_check": {"pixels": len(fd_err), "step": step,
                     "max_abs_err_over_max_grad": float(max(fd_err)),
                     "median_abs_err_over_max_grad": float(np.median(fd_err))},
    }
    return line


def bench_rows(args):
    """The rows beyond the alpha solve, on the 320x240 configuration image:
    estimate_alpha_cf with the reference's default ichol (and this package's
    V-cycle default), estimate_foreground_ml, estimate_foreground_cf,
    estimate_alpha_lkm (radius 10), cf_laplacian at radius 2 and the cutout
    file pipeline. Each is timed on the GPU (numpy in, numpy out, as the
    reference's callers hand them over) and compared with the oracle's
    result, which is timed single-threaded on the host in a subprocess."""
    import tempfile
    import numpy as np
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic

    image, trimap, a_true = synthetic.config1()
    h, w = trimap.shape
    tmp = tempfile.mkdtemp()
    img_png, tri_png = os.path.join(tmp, "image.png"), os.path.join(tmp, "trimap.png")
    pm.save_image(img_png, image)
    pm.save_image(tri_png, trimap)
    img8, tri8 = pm.load_image(img_png, "RGB"), pm.load_image(tri_png, "GRAY")
    rows = {
        "estimate_alpha_cf_ichol": lambda: pm.estimate_alpha_cf(image, trimap,
                                                                preconditioner=pm.ichol),
        "estimate_alpha_cf_vcycle": lambda: pm.estimate_alpha_cf(image, trimap),
        "estimate_foreground_ml": lambda: pm.estimate_foreground_ml(image, a_true),
        "estimate_foreground_cf": lambda: pm.estimate_foreground_cf(image, a_true),
        "estimate_alpha_lkm": lambda: pm.estimate_alpha_lkm(image, trimap),
        "cf_laplacian_radius2": lambda: pm.cf_laplacian(image, radius=2).data,
        "cutout": lambda: (pm.cutout(img_png, tri_png, os.path.join(tmp, "cut.png")),
                           pm.load_image(os.path.join(tmp, "cut.png")))[1],
    }
    out, res = {}, {}
    for name, fn in rows.items():
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = fn()
        torch.cuda.synchronize()
        out[name] = np.asarray(r)
        res[name] = {"ms": round((time.perf_counter() - t0) / args.steps * 1e3, 3),
                     "Mpix_per_s": round(h * w / ((time.perf_counter() - t0) / args.steps) / 1e6,
                                         3)}
    # The oracle on the host, one thread, same inputs.
    opath = os.path.join(tmp, "oracle.npz")
    code = (
        "import sys, time, json, os; sys.path.insert(0, %r)\n"
        "import numpy as np\n"
        "import oracle.cf as ocf, oracle.lkm as olkm\n"
        "from oracle.foreground import estimate_foreground_ml as ofg\n"
        "from oracle.foreground_cf import estimate_foreground_cf as ofcf\n"
        "from oracle.ichol import ichol\n"
        "from oracle.solver import vcycle\n"
        "from PIL import Image\n"
        "from pymatting_amd import synthetic\n"
        "image, trimap, a_true = synthetic.config1()\n"
        "tmp = %r\n"
        "def cut():\n"
        "    im = np.array(Image.open(os.path.join(tmp, 'image.png')).convert('RGB')) / 255.0\n"
        "    tr = np.array(Image.open(os.path.join(tmp, 'trimap.png')).convert('L')) / 255.0\n"
        "    a = ocf.estimate_alpha_cf(im, tr)\n"
        "    F = ofg(im, a)\n"
        "    c = np.concatenate([F, a[:, :, None]], axis=2)\n"
        "    return np.clip(c * 255, 0, 255).astype(np.uint8) / 255.0\n"
        "rows = {'estimate_alpha_cf_ichol': lambda: ocf.estimate_alpha_cf(image, trimap),\n"
        "        'estimate_alpha_cf_vcycle': lambda: ocf.estimate_alpha_cf(image, trimap),\n"
        "        'estimate_foreground_ml': lambda: ofg(image, a_true),\n"
        "        'estimate_foreground_cf': lambda: ofcf(image, a_true),\n"
        "        'estimate_alpha_lkm': lambda: olkm.estimate_alpha_lkm(image, trimap),\n"
        "        'cf_laplacian_radius2': lambda: ocf.cf_laplacian(image, radius=2).data,\n"
        "        'cutout': cut}\n"
        "times, outs = {}, {}\n"
        "for k, f in rows.items():\n"
        "    t = time.perf_counter(); outs[k] = np.asarray(f()); times[k] = time.perf_counter() - t\n"
        "np.savez(%r, **outs)\n"
        "print(json.dumps(times))\n" % (ROOT, tmp, opath))
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               HIP_VISIBLE_DEVICES="", CUDA_VISIBLE_DEVICES="")
    if not args.no_cpu:
        p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                           timeout=900)
        if p.returncode == 0:
            times = json.loads(p.stdout.strip().splitlines()[-1])
            ref = dict(np.load(opath))
            for name in rows:
                d = np.abs(out[name].astype(np.float64) - ref[name].astype(np.float64))
                res[name]["cpu_oracle_ms"] = round(times[name] * 1e3, 1)
                res[name]["cpu_oracle_threads"] = 1
                res[name]["max_abs_diff_vs_oracle"] = float(d.max())
        else:
            res["cpu_oracle_error"] = p.stderr[-400:]
    res["_note"] = ("320x240 configuration image; estimate_alpha_cf_vcycle is this package's "
                    "default preconditioner compared with the oracle's ichol-preconditioned "
                    "alpha (both at rtol 1e-7); the cutout's alpha channel is the ichol-free "
                    "default against the oracle's ichol, 8-bit quantised")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true")
    ap.add_argument("--c2-only", action="store_true",
                    help="the 1920x1080 line alone, without the C3-C5 entries")
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c5", "rows"], default="c2",
                    help="c2: 1920x1080 solve (the metric); c3: 64 images of 256..1024 px, "
                         "batched vs one at a time; c4: 3840x2160 5-level V(2,2) solve and "
                         "float32 foreground; c5: gradient of a 16x512x512 batch")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: sharding over several GPUs is a design, not "
              "part of this build; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.pmc_probe:
        pmc_probe()
        return
    if args.workload in ("c3", "c4", "c5", "rows"):
        print(json.dumps({"c3": bench_c3, "c4": bench_c4, "c5": bench_c5,
                          "rows": bench_rows}[args.workload](args)))
        return

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

    # The configuration: both forms to rtol 1e-7; the mixed form's alpha within
    # max|d alpha| <= 1e-4 of the float64 PCG's (checked below). The mixed
    # form's float64 loop: reliable updates (one float32 CG, float64 solution,
    # float64 residual replacement); restarted refinement beside it. The
    # strict pair runs both forms to rtol 1e-9, where both are within 1e-4 of
    # the converged float64 alpha.
    MIXED = dict(precision="mixed", level0="stencil", rtol=1e-7, mixed="reliable", delta=0.01)
    FP64 = dict(precision="f64", level0="stencil", rtol=1e-7)
    MIXED_STRICT = dict(MIXED, rtol=1e-9)
    FP64_STRICT = dict(FP64, rtol=1e-9)
    MIXED_REFINE = dict(precision="mixed", level0="stencil", rtol=1e-7, inner_rtol=1e-3)
    MIXED_MF = dict(MIXED, level0="matrix_free")
    MIXED_MFR = dict(MIXED, level0="matrix_free_recompute")
    MIXED_G = dict(MIXED, graph=True)

    last = {}

    def step(cfg):
        kw = dict(cfg)
        solver = pm.LambdaSolver(image, trimap, precision=kw.pop("precision"),
                                 level0=kw.pop("level0"))
        alpha = solver.solve(**kw)
        last["solver"] = solver
        return alpha, solver.stats

    def timed(cfg, steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            alpha, st = step(cfg)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        return alpha, st, t

    for _ in range(args.warmup):
        for cfg in (MIXED, FP64, MIXED_STRICT, FP64_STRICT, MIXED_REFINE, MIXED_MF, MIXED_MFR,
                    MIXED_G):
            step(cfg)
    torch.cuda.synchronize()

    alpha_m, st_m, t_mixed = timed(MIXED, args.steps)
    # Roofline kernel: the finest-grid stored-stencil SpMV on this solve's operator.
    op = last["solver"].op
    xs = torch.rand(n, device="cuda", dtype=torch.float32)
    ys = torch.empty_like(xs)
    t_spmv32 = time_spmv(L, op.handle, _ffi.HM_F32, xs, ys)
    xd = xs.double()
    yd = torch.empty_like(xd)
    t_spmv64 = time_spmv(L, op.handle, _ffi.HM_F64, xd, yd)
    # The roofline kernel: the finest grid's pre-smooth + residual +
    # restriction of the value form's own V-cycle (the longest kernel of its
    # iteration), with its prolongation + post-smooth beside it; 50
    # back-to-back launches each, HIP events on the library's stream.
    r_us, s_us = ctypes.c_double(), ctypes.c_double()
    _ffi.check(L.hm_vcycle_time_fine(last["solver"].M.handle, _ffi.ptr(xs), 50, ctypes.byref(r_us),
                                     ctypes.byref(s_us), _ffi.stream()))
    t_restrict, t_smooth = r_us.value * 1e-6, s_us.value * 1e-6
    # Separate pass with per-launch events (they would perturb the timing).
    def profile(cfg):
        """Per-class event timing of one extra step of cfg."""
        return profile_classes(L, lambda: step(cfg))

    prof = profile(MIXED)
    prof64 = profile(FP64)
    prof_mf = profile(MIXED_MF)
    prof_mfr = profile(MIXED_MFR)
    alpha_d, st_d, t_f64 = timed(FP64, args.steps)
    alpha_ms, st_ms, t_ms = timed(MIXED_STRICT, args.steps)
    alpha_ds, st_ds, t_ds = timed(FP64_STRICT, args.steps)
    alpha_f, st_f, t_mf = timed(MIXED_MF, args.steps)
    alpha_fr, st_fr, t_mfr = timed(MIXED_MFR, args.steps)
    alpha_g, st_g, t_g = timed(MIXED_G, args.steps)
    alpha_x, st_x, t_rf = timed(MIXED_REFINE, args.steps)

    # The value form from host buffers (numpy in, numpy out), as the reference's
    # callers hand them over: adds the image/trimap upload and alpha download.
    def host_step():
        kw = dict(MIXED)
        solver = pm.LambdaSolver(image_h, trimap_h, precision=kw.pop("precision"),
                                 level0=kw.pop("level0"))
        return solver.solve(**kw).cpu().numpy()

    def timed_fn(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = fn()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) / args.steps

    alpha_host, t_host = timed_fn(host_step)
    # The reference's own entry point on the same image: estimate_alpha_cf (the
    # unknown-pixel system, V-cycle PCG at its default rtol 1e-7, float64).
    alpha_a1, t_a1 = timed_fn(lambda: pm.estimate_alpha_cf(image, trimap))

    # Converged float64 solution (outside the timed regions).
    conv = pm.LambdaSolver(image, trimap, precision="f64", level0="stencil")
    alpha_c = conv.solve(rtol=1e-11, maxiter=100000)

    def err(a, ref=None):
        d = (a - (alpha_c if ref is None else ref)).abs()
        return float(d.max()), float(d.pow(2).mean().sqrt())

    e_m, e_d, e_f, e_x = err(alpha_m), err(alpha_d), err(alpha_f), err(alpha_x)
    e_ms, e_ds = err(alpha_ms), err(alpha_ds)
    v_m, v_f, v_x = err(alpha_m, alpha_d), err(alpha_f, alpha_d), err(alpha_x, alpha_d)
    e_fr, v_fr = err(alpha_fr), err(alpha_fr, alpha_d)

    bytes_launch = spmv_bytes_per_point(4) * n
    achieved = bytes_launch / t_spmv32 / 1e9
    achieved64 = spmv_bytes_per_point(8) * n / t_spmv64 / 1e9
    restrict_bytes = bytes_per_point("dia_restrict", 4) * n
    smooth_bytes = bytes_per_point("dia_smooth", 4) * n
    achieved_r = restrict_bytes / t_restrict / 1e9
    achieved_s = smooth_bytes / t_smooth / 1e9
    def summarise(pr):
        """Per-class launches, time and algorithmic GB/s (events around each
        launch add a few microseconds to each), and the step's bytes."""
        table, total = {}, 0.0
        for (k, p), v in sorted(pr.items(), key=lambda kv: -kv[1][0]):
            ms, cnt, pts, s = v
            entry = {"ms_total": round(ms, 3), "launches": cnt,
                     "avg_us_with_events": round(ms / cnt * 1e3, 2)}
            bpp = bytes_per_point(k, s)
            if bpp is not None:
                gbps = bpp * pts / (ms * 1e-3) / 1e9
                entry["bytes_per_launch"] = int(bpp * pts / cnt)
                entry["GBps_with_events"] = round(gbps, 1)
                entry["frac_of_hbm_peak"] = round(gbps / HBM_PEAK_GBS, 4)
                total += bpp * pts
            table["%s_%s" % (k, "f32" if p else "f64")] = entry
        return table, total

    breakdown, step_bytes = summarise(prof)
    breakdown64, step_bytes64 = summarise(prof64)
    # The three forms of the finest-grid operator in the value form: stored
    # 13-slot stencil, matrix-free with stored window planes, matrix-free with
    # the windows recomputed per pass.
    level0_forms = {
        "stored_stencil": fine_passes(prof, "dia"),
        "matrix_free_stored_windows": fine_passes(prof_mf, "cf"),
        "matrix_free_recomputed_windows": fine_passes(prof_mfr, "cfrec"),
    }
    # The solve's own kernels (fine grid: the fused SpMV with the direction
    # update, the pre-smooth + residual + restriction, the prolongation +
    # post-smooth; the first Galerkin grid; the coarser grids) as
    # algorithmic bytes per launch over the mean event-timed launch.
    solve_kernels = {k: {q: breakdown[k][q] for q in ("bytes_per_launch", "avg_us_with_events",
                                                       "GBps_with_events", "frac_of_hbm_peak")}
                     for k in ("dia_apply_f32", "dia_restrict_f32", "dia_smooth_f32",
                               "l1_restrict_f32", "l1_smooth_f32", "coarse_restrict_f32",
                               "coarse_smooth_f32", "pair_up_f32") if k in breakdown}

    oracle_vs = None
    if args.no_cpu:
        cpu = None
    else:
        import tempfile
        fd, apath = tempfile.mkstemp(suffix=".npy")
        os.close(fd)
        cpu = cpu_baseline(apath)
        try:
            a_or = torch.from_numpy(np.load(apath)).cuda()
            oracle_vs = {"mixed": err(alpha_m, a_or), "fp64": err(alpha_d, a_or)}
        except (OSError, ValueError):
            oracle_vs = None
        finally:
            os.unlink(apath)

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
                               "V(1,1) vcycle PCG, rtol 1e-7); mixed float32 CG with float64 "
                               "solution and residual replacement (reliable updates) against the "
                               "float64 PCG at the same rtol, max|dalpha| <= 1e-4; stored 13-slot "
                               "stencil",
                   "height": h, "width": w,
                   "unknown_fraction": round(float((trimap_h == 0.5).mean()), 4),
                   "rtol": MIXED["rtol"], "mixed_delta": MIXED["delta"]},
        "mixed": {"iterations": st_m["iterations"],
                  "within_1e-4_of_fp64": bool(v_m[0] <= 1e-4),
                  "max_abs_dalpha_vs_fp64": v_m[0], "rms_dalpha_vs_fp64": v_m[1],
                  "max_abs_dalpha_vs_converged": e_m[0], "rms_dalpha_vs_converged": e_m[1]},
        "fp64": {"value": round(n / t_f64 / 1e6, 3), "unit": "Mpix/s",
                 "ms_per_step": round(t_f64 * 1e3, 3), "rtol": FP64["rtol"],
                 "iterations": st_d["iterations"], "max_abs_dalpha_vs_converged": e_d[0],
                 "rms_dalpha_vs_converged": e_d[1]},
        "vs_cpu_oracle_rtol_1e-7": None if oracle_vs is None else {
            k: {"max_abs_dalpha": v[0], "rms_dalpha": v[1]} for k, v in oracle_vs.items()},
        "strict_vs_converged": {
            "rtol": MIXED_STRICT["rtol"],
            "mixed": {"value": round(n / t_ms / 1e6, 3), "unit": "Mpix/s",
                      "ms_per_step": round(t_ms * 1e3, 3), "iterations": st_ms["iterations"],
                      "max_abs_dalpha": e_ms[0], "rms_dalpha": e_ms[1]},
            "fp64": {"value": round(n / t_ds / 1e6, 3), "unit": "Mpix/s",
                     "ms_per_step": round(t_ds * 1e3, 3), "iterations": st_ds["iterations"],
                     "max_abs_dalpha": e_ds[0], "rms_dalpha": e_ds[1]}},
        "mixed_matrix_free": {"value": round(n / t_mf / 1e6, 3), "unit": "Mpix/s",
                              "iterations": st_f["iterations"],
                              "max_abs_dalpha_vs_fp64": v_f[0],
                              "max_abs_dalpha_vs_converged": e_f[0]},
        "mixed_matrix_free_recompute": {"value": round(n / t_mfr / 1e6, 3), "unit": "Mpix/s",
                                        "iterations": st_fr["iterations"],
                                        "max_abs_dalpha_vs_fp64": v_fr[0],
                                        "max_abs_dalpha_vs_converged": e_fr[0]},
        "fine_passes_by_level0_form_f32": level0_forms,
        # The value form with its iterations without a replacement replayed
        # from HIP graphs instead of launched directly.
        "mixed_graph_replay": {"value": round(n / t_g / 1e6, 3), "unit": "Mpix/s",
                               "ms_per_step": round(t_g * 1e3, 3),
                               "iterations": st_g["iterations"],
                               "identical_to_direct": bool(torch.equal(alpha_g, alpha_m))},
        "mixed_host_buffers": {"value": round(n / t_host / 1e6, 3), "unit": "Mpix/s",
                               "ms_per_step": round(t_host * 1e3, 3),
                               "max_abs_dalpha_vs_device": float(np.abs(
                                   alpha_host - alpha_m.cpu().numpy()).max())},
        "estimate_alpha_cf_fp64": {"value": round(n / t_a1 / 1e6, 3), "unit": "Mpix/s",
                                   "ms_per_step": round(t_a1 * 1e3, 3), "rtol": 1e-7},
        "mixed_restarted_refinement": {"value": round(n / t_rf / 1e6, 3), "unit": "Mpix/s",
                                       "rtol": MIXED_REFINE["rtol"],
                                       "inner_rtol": MIXED_REFINE["inner_rtol"],
                                       "iterations": st_x["iterations"],
                                       "max_abs_dalpha_vs_fp64": v_x[0],
                                       "max_abs_dalpha_vs_converged": e_x[0],
                                       "rms_dalpha_vs_converged": e_x[1]},
        "roofline": {"bound": "hbm",
                     "kernel": "tile_kernel<float, DiaOp<float, false>, InJac0<float>, "
                               "OutRestrict<float, 0>, 0> (finest-grid pre-smooth + residual + "
                               "restriction of the value form's V-cycle, 1920x1080)",
                     "achieved": round(achieved_r, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_r / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC_RESTRICT_F32,
                     "bytes_per_launch": int(restrict_bytes),
                     "avg_launch_us": round(t_restrict * 1e6, 2),
                     "fine_smooth": {"kernel": "tile_kernel<float, DiaOp<float, false>, "
                                               "InProlong<float, InJac0<float> >, "
                                               "OutSmooth<float, 0>, 0>",
                                     "achieved": round(achieved_s, 1),
                                     "frac": round(achieved_s / HBM_PEAK_GBS, 4),
                                     "traffic": TRAFFIC_SMOOTH_F32,
                                     "bytes_per_launch": int(smooth_bytes),
                                     "avg_launch_us": round(t_smooth * 1e6, 2)},
                     "proxy_spmv": {"kernel": "tile_kernel<float, DiaOp<float, false>, "
                                              "InVec<float>, OutApply<float, 0>, 0> (y = A x on "
                                              "the solve's stored operator)",
                                    "achieved": round(achieved, 1),
                                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                                    "traffic": TRAFFIC_SPMV_F32,
                                    "bytes_per_launch": int(bytes_launch),
                                    "avg_launch_us": round(t_spmv32 * 1e6, 2),
                                    "f64": {"achieved": round(achieved64, 1),
                                            "frac": round(achieved64 / HBM_PEAK_GBS, 4),
                                            "avg_launch_us": round(t_spmv64 * 1e6, 2)}}},
        "solve_kernels_mixed": solve_kernels,
        "kernel_breakdown_timed_mixed": breakdown,
        "kernel_breakdown_timed_fp64": breakdown64,
        # Algorithmic bytes of one step (the classes above; the LDS tail's
        # launches, with the x update they carry, are not counted) over the
        # step's time, for both forms.
        "step_bytes_model": {"GB": round(step_bytes / 1e9, 3),
                             "GBps_over_step": round(step_bytes / t_mixed / 1e9, 1),
                             "frac_of_hbm_peak": round(step_bytes / t_mixed / 1e9
                                                       / HBM_PEAK_GBS, 4)},
        "step_bytes_model_fp64": {"GB": round(step_bytes64 / 1e9, 3),
                                  "GBps_over_step": round(step_bytes64 / t_f64 / 1e9, 1),
                                  "frac_of_hbm_peak": round(step_bytes64 / t_f64 / 1e9
                                                            / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
    }
    if not args.c2_only:
        # The other configurations' lines, measured in the same run: C3 (64
        # batched images; the refinement form only under --workload c3), C4
        # (3840x2160 and its float32 foreground) and C5 (16x512x512
        # gradient), each with its accuracy figures.
        del last["solver"], op
        torch.cuda.empty_cache()
        line["c3"] = _sub_line(bench_c3(args, ("f64", "mixed", "f64_rtol_1e-7",
                                               "mixed_rtol_1e-7")))
        torch.cuda.empty_cache()
        line["c4"] = _sub_line(bench_c4(args))
        torch.cuda.empty_cache()
        line["c5"] = _sub_line(bench_c5(args))
        torch.cuda.empty_cache()
        line["other_rows_320x240"] = bench_rows(args)
    print(json.dumps(line))


def _sub_line(line):
    """A configuration's line as an entry of the main line (without the keys
    that repeat the run's own)."""
    return {k: v for k, v in line.items()
            if k not in ("n_gpus", "steps", "warmup", "scaling", "vs_baseline", "data")}


if __name__ == "__main__":
    main()
