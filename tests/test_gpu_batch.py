"""Batched PCG (BatchSolver, hm_batch_*) against one-image solves: same
solutions, per-image convergence, graph replay identical to direct launches,
every tile shape, and the mixed form's accuracy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(64, 80), (96, 64), (128, 128), (80, 112), (72, 72), (136, 48)]


def _images():
    from pymatting_amd import synthetic
    out = []
    for k, (h, w) in enumerate(SIZES):
        image, trimap, _ = synthetic.make_image(h, w, 3100 + k, strands=5, erode=3 + k % 3)
        out.append((image, trimap))
    return out


def test_batch_matches_single_image_solves():
    import pymatting_amd as pm
    data = _images()
    bs = pm.BatchSolver([d[0] for d in data], [d[1] for d in data], precision="f64",
                        graph_iters=0)
    alphas = bs.solve(rtol=1e-10)
    for (image, trimap), a, st in zip(data, alphas, bs.stats):
        s = pm.LambdaSolver(image, trimap, precision="f64", level0="stencil")
        ar = s.solve(rtol=1e-10)
        assert float((a - ar).abs().max()) < 1e-7
        assert st["converged"]
        assert abs(st["iterations"] - s.stats["iterations"]) <= 1


def test_graph_replay_equals_direct_launches_and_shapes_agree():
    import pymatting_amd as pm
    data = _images()
    bs = pm.BatchSolver([d[0] for d in data], [d[1] for d in data], precision="f64",
                        graph_iters=0)
    x0 = [x.clone() for x in bs.solve_x(rtol=1e-9)]
    it0 = [s["iterations"] for s in bs.stats]
    for g in (2, 4, 6):
        bs.graph_iters = g
        for rep in range(2):  # replays of a fresh graph, twice
            xg = bs.solve_x(rtol=1e-9)
            assert [s["iterations"] for s in bs.stats] == it0
            for a, b in zip(x0, xg):
                assert bool(torch_equal(a, b))
    bs.graph_iters = 4
    for s0 in (0, 1):
        for sc in (0, 1):
            bs.level0_shape, bs.coarse_shape = s0, sc
            xs = bs.solve_x(rtol=1e-9)
            for a, b in zip(x0, xs):
                assert float((a - b).abs().max()) < 1e-7


def torch_equal(a, b):
    import torch
    return torch.equal(a, b)


def test_batch_mixed_accuracy_and_masks():
    import pymatting_amd as pm
    data = _images()
    ref = [pm.LambdaSolver(im, tr, precision="f64", level0="stencil").solve(rtol=1e-11)
           for im, tr in data]
    bs = pm.BatchSolver([d[0] for d in data], [d[1] for d in data], precision="mixed",
                        graph_iters=4)
    alphas = bs.solve(rtol=1e-8, inner_rtol=1e-3)
    for a, r, st in zip(alphas, ref, bs.stats):
        assert st["converged"]
        assert float((a - r).abs().max()) < 1e-4
    # Iteration counts differ per image: each stops at its own tolerance.
    assert len({st["iterations"] for st in bs.stats}) > 1


def test_batch_autotune_and_errors():
    import pymatting_amd as pm
    data = _images()[:3]
    bs = pm.BatchSolver([d[0] for d in data], [d[1] for d in data], precision="f64")
    choice = bs.autotune(rtol=1e-7)
    assert choice[0] in (0, 1) and choice[1] in (0, 1) and choice[2] in (0, 2, 4, 8, 16)
    assert bs.autotune(rtol=1e-7) == choice  # cached
    bs.solve(rtol=1e-7)
    # Reliable updates tune the shapes and graph replay (8) against direct
    # launches (0), cached per form.
    m = pm.BatchSolver([d[0] for d in data], [d[1] for d in data], precision="mixed")
    rel = m.autotune(rtol=1e-9, mixed="reliable")
    assert rel[0] in (0, 1) and rel[1] in (0, 1) and rel[2] in (0, 8)
    assert m.autotune(rtol=1e-9, mixed="reliable") == rel
    for st in m.stats:
        assert st["converged"]
    with pytest.raises(ValueError):
        pm.BatchSolver([data[0][0]], [data[0][1], data[1][1]])
    # V-cycles with different sweep counts cannot share a batch.
    a = pm.BatchSolver([d[0] for d in data], [d[1] for d in data])
    a.solvers[1] = pm.LambdaSolver(data[1][0], data[1][1], level0="stencil", num_pre_iter=2,
                                   num_post_iter=2)
    with pytest.raises(ValueError, match="same V-cycle"):
        a.solve()


def _true_rel(solver, x, b):
    import torch
    r = b - solver.op.apply(x)
    return float(torch.linalg.norm(r) / torch.linalg.norm(b))


def test_batch_reliable_updates():
    """Batched mixed solve with reliable updates: every image reaches a
    float64 residual below rtol in about the float64 iteration count, and its
    alpha is within 1e-4 of the converged one."""
    import pymatting_amd as pm
    data = _images()
    ref = [pm.LambdaSolver(im, tr, precision="f64", level0="stencil") for im, tr in data]
    conv = [s.solve(rtol=1e-11) for s in ref]
    its64 = []
    for s in ref:
        s.solve_system(s.b, rtol=1e-9)
        its64.append(s.stats["iterations"])
    bs = pm.BatchSolver([d[0] for d in data], [d[1] for d in data], precision="mixed")
    xs = bs.solve_x(rtol=1e-9, mixed="reliable")
    for x, s, c, st, it in zip(xs, bs.solvers, conv, bs.stats, its64):
        print(st, it, flush=True)
        assert st["converged"]
        assert _true_rel(s, x, s.b) < 1e-9
        assert abs(st["iterations"] - it) <= 2
        assert float((x.clamp(0, 1).reshape(c.shape) - c).abs().max()) < 1e-4


def test_batch_reliable_graph_replay_equals_direct_launches():
    """Reliable updates with each iteration's fixed parts replayed from HIP
    graphs (residual replacements, state read-backs and restarts launched
    directly between the replays) give bit-identical solutions and iteration
    counts to direct launches, on two solves in a row of each kind, also when
    images restart."""
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    data = _images()
    im1, tr1, _ = synthetic.config1()
    cases = [([d[0] for d in data], [d[1] for d in data], dict(rtol=1e-9)),
             ([im1, data[2][0]], [tr1, data[2][1]], dict(rtol=1e-11, delta=1e-12, maxiter=2000))]
    for images, trimaps, kw in cases:
        bs = pm.BatchSolver(images, trimaps, precision="mixed")
        bs.graph_iters = 0
        x0 = [x.clone() for x in bs.solve_x(mixed="reliable", **kw)]
        it0 = [st["iterations"] for st in bs.stats]
        assert all(st["converged"] for st in bs.stats)
        for rep in range(2):
            bs.graph_iters = 8
            xg = bs.solve_x(mixed="reliable", **kw)
            print(kw, rep, [st["iterations"] for st in bs.stats], it0, flush=True)
            assert [st["iterations"] for st in bs.stats] == it0
            for a, b in zip(x0, xg):
                assert torch.equal(a, b)
            bs.graph_iters = 0
            xd = bs.solve_x(mixed="reliable", **kw)
            for a, b in zip(x0, xd):
                assert torch.equal(a, b)


def test_batch_reliable_restart_and_zero_rhs():
    """Fast-converging systems (random colours, 60% known pixels) converge
    between read-backs: the float64 check fails on some images and they are
    restarted without disturbing the others; a zero right-hand side is
    solved by x = 0."""
    import torch
    import pymatting_amd as pm
    rng = np.random.default_rng(7)
    images, trimaps = [], []
    for h, w in ((97, 131), (64, 64), (120, 90)):
        images.append(rng.random((h, w, 3)))
        trimaps.append(rng.choice([0.0, 0.5, 1.0], size=(h, w), p=[0.3, 0.4, 0.3]))
    bs = pm.BatchSolver(images, trimaps, precision="mixed")
    rhs = [s.b.clone() for s in bs.solvers]
    rhs[1].zero_()
    for rtol in (1e-9, 1e-11):
        xs = bs.solve_x(rtol=rtol, mixed="reliable", rhs=rhs)
        for k, (x, s, st) in enumerate(zip(xs, bs.solvers, bs.stats)):
            print(rtol, k, st, flush=True)
            assert st["converged"]
            if k == 1:
                assert float(x.abs().max()) == 0.0
                continue
            f64 = pm.LambdaSolver(images[k], trimaps[k], precision="f64", level0="stencil")
            f64.solve_system(f64.b, rtol=rtol)
            assert _true_rel(s, x, s.b) < rtol
            assert st["iterations"] <= f64.stats["iterations"] + 2
    with pytest.raises(ValueError):
        bs.solve_x(mixed="other")
    # Replacements that are never due (delta 1e-12) on systems where the
    # float32 recursion stalls above rtol 1e-11: the float64 check fails, the
    # images restart one by one without disturbing the others, and each ends
    # below rtol in float64.
    from pymatting_amd import synthetic
    im1, tr1, _ = synthetic.config1()
    im2, tr2, _ = synthetic.make_image(200, 260, 77, strands=8, erode=6)
    hard = pm.BatchSolver([im1, im2], [tr1, tr2], precision="mixed")
    xs = hard.solve_x(rtol=1e-11, mixed="reliable", delta=1e-12, maxiter=2000)
    for x, s, st in zip(xs, hard.solvers, hard.stats):
        print("restarts", st, flush=True)
        assert st["converged"]
        assert _true_rel(s, x, s.b) < 1e-11


def test_time_sync_probe():
    """hm_time_sync: a kernel boundary and both grid barriers are timed (all
    barriers complete), and grids larger than one workgroup per compute unit
    are refused."""
    import ctypes
    from pymatting_amd import _ffi
    L = _ffi.lib()
    lu, cu, au, ok = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    for G in (8, 64):
        _ffi.check(L.hm_time_sync(G, 200, ctypes.byref(lu), ctypes.byref(cu), ctypes.byref(au),
                                  ctypes.byref(ok), _ffi.stream()))
        print(G, lu.value, cu.value, au.value, ok.value, flush=True)
        assert ok.value == 1
        assert 0.0 < lu.value < 1e3 and 0.0 < cu.value < 1e3 and 0.0 < au.value < 1e3
    assert L.hm_time_sync(100000, 10, ctypes.byref(lu), ctypes.byref(cu), ctypes.byref(au),
                          ctypes.byref(ok), _ffi.stream()) == _ffi.HM_ERR_ARG
    assert L.hm_time_sync(8, 0, ctypes.byref(lu), ctypes.byref(cu), ctypes.byref(au),
                          ctypes.byref(ok), _ffi.stream()) == _ffi.HM_ERR_ARG
