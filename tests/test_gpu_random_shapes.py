"""Parity on seeded random shapes (3..220 px a side, some 3..7 px thin) and
random trimaps, against the oracle: cf_laplacian bit-exact; the float64 and
mixed lambda solves and estimate_alpha_cf on the converged solutions; the
float32 foreground bit-exact; and a ragged batch of eight images in every
batched form."""
import warnings

import numpy as np
import pytest

import oracle.cf as ocf
from oracle.foreground import estimate_foreground_ml as ofg

pytestmark = pytest.mark.gpu


def _case(rng):
    h = int(rng.integers(3, 220))
    w = int(rng.integers(3, 220))
    if rng.random() < 0.2:
        h = int(rng.integers(3, 8))
    if rng.random() < 0.2:
        w = int(rng.integers(3, 8))
    image = rng.random((h, w, 3))
    if rng.random() < 0.3:  # smooth colours: near-singular window covariances
        yy, xx = np.mgrid[0:h, 0:w]
        image = np.stack([0.5 + 0.4 * np.sin(xx / 17.0 + c) * np.cos(yy / 13.0)
                          for c in range(3)], -1)
    trimap = rng.choice([0.0, 0.5, 1.0], size=(h, w), p=[0.35, 0.3, 0.35])
    trimap.flat[0], trimap.flat[-1] = 0.0, 1.0
    return image, trimap


@pytest.mark.parametrize("seed", range(12))
def test_random_shape(seed):
    import pymatting_amd as pm
    rng = np.random.default_rng(1000 + seed)
    image, trimap = _case(rng)
    print("seed", seed, "shape", image.shape, flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # the oracle's ichol may report a shift
        L, R = pm.cf_laplacian(image), ocf.cf_laplacian(image)
        assert np.array_equal(L.data, R.data)
        ref = ocf.solve_alpha_lambda(image, trimap, rtol=1e-11)[0]
        a64 = pm.LambdaSolver(image, trimap, precision="f64", level0="stencil").solve(rtol=1e-11)
        assert np.abs(a64.cpu().numpy() - ref).max() < 1e-6
        m = pm.LambdaSolver(image, trimap, precision="mixed", level0="stencil")
        am = m.solve(rtol=1e-10, mixed="reliable")
        assert np.abs(am.cpu().numpy() - ref).max() < 1e-5
        a1 = pm.estimate_alpha_cf(image, trimap, cg_kwargs={"rtol": 1e-12})
        a1r = ocf.estimate_alpha_cf(image, trimap, cg_kwargs={"rtol": 1e-12})
        assert np.abs(a1 - a1r).max() < 1e-6
    alpha = np.clip(rng.random(trimap.shape) * 1.4 - 0.2, 0, 1)
    F = pm.estimate_foreground_ml(image, alpha)
    assert np.array_equal(F, ofg(image, alpha, mean_mode="f64"))


def test_random_ragged_batch():
    import torch
    import pymatting_amd as pm
    rng = np.random.default_rng(77)
    ims, tris, refs = [], [], []
    for _ in range(8):
        h, w = int(rng.integers(16, 200)), int(rng.integers(16, 200))
        image = rng.random((h, w, 3))
        trimap = rng.choice([0.0, 0.5, 1.0], size=(h, w), p=[0.35, 0.3, 0.35])
        trimap.flat[0], trimap.flat[-1] = 0.0, 1.0
        ims.append(torch.from_numpy(image).cuda())
        tris.append(torch.from_numpy(trimap).cuda())
        refs.append(ocf.solve_alpha_lambda(image, trimap, rtol=1e-11)[0])
    for prec, kw, tol in (("f64", dict(rtol=1e-11), 1e-6),
                          ("mixed", dict(rtol=1e-10, mixed="reliable"), 1e-5),
                          ("mixed", dict(rtol=1e-10, inner_rtol=1e-3), 1e-5)):
        alphas = pm.BatchSolver(ims, tris, precision=prec).solve(**kw)
        for a, r in zip(alphas, refs):
            assert np.abs(a.cpu().numpy() - r).max() < tol, (prec, kw)
