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
    with pytest.raises(ValueError):
        pm.BatchSolver([data[0][0]], [data[0][1], data[1][1]])
    # V-cycles with different sweep counts cannot share a batch.
    a = pm.BatchSolver([d[0] for d in data], [d[1] for d in data])
    a.solvers[1] = pm.LambdaSolver(data[1][0], data[1][1], level0="stencil", num_pre_iter=2,
                                   num_post_iter=2)
    with pytest.raises(ValueError, match="same V-cycle"):
        a.solve()
