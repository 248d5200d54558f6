"""The single-image value form (reliable updates) with its iterations
without a replacement replayed from HIP graphs (hm_cg_solve_reliable,
graph = 1) against direct launches: bit-identical solutions and iteration
counts on the stored and both matrix-free level-0 operators, for several
read-back intervals, on two solves in a row of each kind (the graphs are
recorded per solve and replayed many times, with residual replacements,
read-backs and host work between the replays)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("level0", ["stencil", "matrix_free", "matrix_free_recompute"])
def test_graph_replay_equals_direct_launches(level0):
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    image, trimap, _ = synthetic.config1()
    img, tri = torch.from_numpy(image).cuda(), torch.from_numpy(trimap).cuda()
    s = pm.LambdaSolver(img, tri, precision="mixed", level0=level0)
    for kw in (dict(rtol=1e-7), dict(rtol=1e-10, check_every=3), dict(rtol=1e-9, delta=0.3)):
        x0 = s.solve_system(s.b, mixed="reliable", **kw).clone()
        it0 = s.stats["iterations"]
        for rep in range(2):
            xg = s.solve_system(s.b, mixed="reliable", graph=True, **kw)
            print(level0, kw, rep, s.stats["iterations"], it0, flush=True)
            assert s.stats["iterations"] == it0
            assert torch.equal(xg, x0)
            xd = s.solve_system(s.b, mixed="reliable", **kw)
            assert torch.equal(xd, x0)


def test_graph_argument_checked():
    import ctypes
    from pymatting_amd import _ffi
    st = _ffi.CgStats()
    assert _ffi.lib().hm_cg_solve_reliable(None, None, None, None, None, 1e-7, 0.01, 10, 8, 2,
                                           ctypes.byref(st), None) == _ffi.HM_ERR_ARG
