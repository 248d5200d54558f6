"""Several coarse V-cycle levels per launch (csrc/pair.h: ascent chains,
descent pairs) against one level per launch: the cycle, and PCG with it,
must agree bit for bit, in float64 and float32, on grids with odd and thin
coarse levels. The one-level path is itself pinned against the oracle's
vcycle (test_gpu_solver.py)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _solver(h, w, seed, precision):
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    image, trimap, _ = synthetic.make_image(h, w, seed=seed, erode=4)
    return pm.LambdaSolver(torch.from_numpy(image).cuda(), torch.from_numpy(trimap).cuda(),
                           precision=precision, level0="stencil")


@pytest.mark.parametrize("precision", ["f64", "mixed"])
@pytest.mark.parametrize("shape", [(1080, 1920), (333, 517), (129, 700), (700, 66), (190, 190)])
def test_pairs_equal_single_levels(shape, precision):
    h, w = shape
    print("pairs vs single levels", shape, precision, flush=True)
    s = _solver(h, w, 7, precision)
    M = s.M
    dt = torch.float64 if precision == "f64" else torch.float32
    g = torch.Generator().manual_seed(h * 7 + w)
    r = torch.rand(h * w, generator=g, dtype=torch.float64).to(dt).cuda()
    z_single = M.set_pairs(0).apply(r).clone()
    torch.cuda.synchronize()
    assert torch.isfinite(z_single).all()
    for mode in (1, 2, 3):
        assert torch.equal(M.set_pairs(mode).apply(r), z_single), mode
    # A second application (the hierarchy's buffers hold the previous cycle).
    r2 = torch.rand(h * w, generator=g, dtype=torch.float64).to(dt).cuda()
    z2 = M.set_pairs(0).apply(r2).clone()
    assert torch.equal(M.set_pairs(3).apply(r2), z2)
    assert torch.equal(M.set_pairs(1).apply(r2), z2)
    if shape == (1080, 1920):  # grids 480x270 .. 60x33 run together
        down, up = _pair_launches(lambda: M.set_pairs(1).apply(r))
        assert down == 0 and 1 <= up <= 2
        down, up = _pair_launches(lambda: M.set_pairs(3).apply(r))
        assert down == 2 and 1 <= up <= 2
        assert _pair_launches(lambda: M.set_pairs(0).apply(r)) == (0, 0)
    M.set_pairs(1)


def test_set_pairs_rejects_unknown_modes():
    from pymatting_amd import _ffi
    s = _solver(64, 64, 1, "f64")
    with pytest.raises(Exception, match="pairs must be 0..3"):
        s.M.set_pairs(4)


def _pair_launches(fn):
    import ctypes
    import bench
    from pymatting_amd import _ffi
    L = _ffi.lib()
    L.hm_prof_set(1)
    torch.cuda.synchronize()
    fn()
    torch.cuda.synchronize()
    out = []
    for kind in ("pair_down", "pair_up"):
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
def test_pcg_with_pairs_equals_single_levels(precision):
    kw = dict(rtol=1e-7)
    if precision == "mixed":
        kw.update(mixed="reliable")
    s = _solver(270, 480, 3, precision)
    s.M.set_pairs(0)
    a_single = s.solve(**kw).clone()
    it_single = s.stats["iterations"]
    for mode in (1, 3):
        s.M.set_pairs(mode)
        a_pairs = s.solve(**kw)
        assert s.stats["iterations"] == it_single
        assert torch.equal(a_pairs, a_single)
