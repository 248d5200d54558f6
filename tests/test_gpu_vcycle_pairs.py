"""The V-cycle's launch structures (hm_vcycle_set_launches) against one grid
per launch. Ascent chains and descent pairs (csrc/pair.h) recompute
intermediate grids on halos with the same arithmetic, so the cycle, and PCG
with it, must agree bit for bit, in float64 and float32, on grids with odd
and thin coarse levels. The tail applied as its precomputed matrix
(csrc/vcycle.h) changes the cycle by rounding only. The one-grid-per-launch
path is itself pinned against the oracle's vcycle (test_gpu_solver.py)."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu

CHAIN, PAIR_DOWN, TAIL_MATRIX = 1, 2, 4


def _solver(h, w, seed, precision):
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    image, trimap, _ = synthetic.make_image(h, w, seed=seed, erode=4)
    return pm.LambdaSolver(torch.from_numpy(image).cuda(), torch.from_numpy(trimap).cuda(),
                           precision=precision, level0="stencil")


def _launches(fn):
    """(descent pairs, ascent chains, tail launches) that fn() made."""
    import bench
    from pymatting_amd import _ffi
    L = _ffi.lib()
    L.hm_prof_set(1)
    torch.cuda.synchronize()
    fn()
    torch.cuda.synchronize()
    out = []
    for kind in ("pair_down", "pair_up", "tail"):
        total = 0
        for prec in (0, 1):
            ms, cnt, pts = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
            L.hm_prof_query(bench.KINDS.index(kind), prec, ctypes.byref(ms), ctypes.byref(cnt),
                            ctypes.byref(pts))
            total += cnt.value
        out.append(total)
    L.hm_prof_set(0)
    return tuple(out)


@pytest.mark.parametrize("precision", ["f64", "mixed"])
@pytest.mark.parametrize("shape", [(1080, 1920), (333, 517), (129, 700), (700, 66), (190, 190)])
def test_launch_structures_agree(shape, precision):
    h, w = shape
    print("launch structures", shape, precision, flush=True)
    s = _solver(h, w, 7, precision)
    M = s.M
    dt = torch.float64 if precision == "f64" else torch.float32
    g = torch.Generator().manual_seed(h * 7 + w)
    r = torch.rand(h * w, generator=g, dtype=torch.float64).to(dt).cuda()
    r2 = torch.rand(h * w, generator=g, dtype=torch.float64).to(dt).cuda()
    z_kernel = None
    for tail in (0, TAIL_MATRIX):
        z_single = M.set_launches(tail).apply(r).clone()
        z2 = M.set_launches(tail).apply(r2).clone()
        torch.cuda.synchronize()
        assert torch.isfinite(z_single).all()
        for mode in (CHAIN, PAIR_DOWN, CHAIN | PAIR_DOWN):
            assert torch.equal(M.set_launches(mode | tail).apply(r), z_single), (mode, tail)
            # Again: the hierarchy's buffers hold the previous cycle's values.
            assert torch.equal(M.set_launches(mode | tail).apply(r2), z2), (mode, tail)
        if tail == 0:
            z_kernel = z_single
    # The tail as a matrix: the same map up to the rounding of its sums.
    z_matrix = M.set_launches(TAIL_MATRIX).apply(r)
    rel = float((z_matrix - z_kernel).abs().max() / z_kernel.abs().max())
    assert rel < (1e-12 if precision == "f64" else 1e-5), rel
    if shape == (1080, 1920):  # grids 480x270 .. 60x33 run in one ascent launch
        assert _launches(lambda: M.set_launches(CHAIN | TAIL_MATRIX).apply(r)) == (0, 1, 1)
        assert _launches(lambda: M.set_launches(CHAIN | PAIR_DOWN).apply(r)) == (2, 1, 1)
        assert _launches(lambda: M.set_launches(0).apply(r)) == (0, 0, 1)
    M.set_launches(CHAIN | TAIL_MATRIX)


def test_set_launches_rejects_unknown_modes():
    s = _solver(64, 64, 1, "f64")
    with pytest.raises(ValueError, match="launch structure must be 0..7"):
        s.M.set_launches(8)


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_pcg_with_launch_structures(precision):
    kw = dict(rtol=1e-7)
    if precision == "mixed":
        kw.update(mixed="reliable")
    s = _solver(270, 480, 3, precision)
    s.M.set_launches(0)
    a_single = s.solve(**kw).clone()
    it_single = s.stats["iterations"]
    for mode in (CHAIN, CHAIN | PAIR_DOWN):
        s.M.set_launches(mode)
        a_pairs = s.solve(**kw)
        assert s.stats["iterations"] == it_single
        assert torch.equal(a_pairs, a_single)
    s.M.set_launches(CHAIN | TAIL_MATRIX)
    a_matrix = s.solve(**kw)
    assert abs(s.stats["iterations"] - it_single) <= 1
    assert float((a_matrix - a_single).abs().max()) < (1e-9 if precision == "f64" else 1e-5)


def test_unaligned_solution_vector():
    """The CG updates stream four consecutive elements per access when the
    vectors allow it and fall back to single elements otherwise: a solution
    vector at an odd float64 offset gives the same solve up to the order of
    the residual-norm sums."""
    from pymatting_amd import _ffi
    s = _solver(270, 480, 5, "mixed")
    n = s.h * s.w
    L = _ffi.lib()
    out = []
    for offset in (0, 1):
        buf = torch.zeros(n + 4, dtype=torch.float64, device="cuda")
        x = buf[offset:offset + n]
        stats = _ffi.CgStats()
        _ffi.check(L.hm_cg_solve_reliable(None, s.op.handle, s.M.handle, _ffi.ptr(s.b),
                                          _ffi.ptr(x), 1e-7, 0.01, 10000, 8, 0,
                                          ctypes.byref(stats), _ffi.stream()))
        out.append((x.clone(), stats.iterations))
    (x0, it0), (x1, it1) = out
    assert abs(it0 - it1) <= 1
    assert float((x0 - x1).abs().max()) < 1e-6 * float(x0.abs().max())
