"""GPU parity of lkm_laplacian / estimate_alpha_lkm
(pymatting/laplacian/lkm_laplacian.py, pymatting/alpha/estimate_alpha_lkm.py)
and of cutout (pymatting/cutout/cutout.py).

- The device operator and diagonal against the oracle's restatement (1e-10
  of the largest magnitude) for radii 1 to 10, and, since the operator is
  the closed-form Laplacian of the same radius, against cf_laplacian.
- The reference's tests/test_lkm.py known answer on the device: lemur at
  0.125, radius 2, Jacobi CG on the large-kernel and on the stored radius-2
  system agree within 1e-5 in iteration counts within 2.
- estimate_alpha_lkm within 1e-6 of the oracle's at a tight tolerance.
- cutout on the 32x32 test image: tests/test_simple_api.py's bounds.
"""
import numpy as np
import pytest

from conftest import load_golden
import oracle.cf as ocf
import oracle.lkm as olkm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("r", [1, 2, 5, 10])
def test_operator_and_diagonal_match_oracle(r):
    import pymatting_amd as pm
    rng = np.random.default_rng(r)
    h, w = 2 * r + 23, 2 * r + 31
    image = rng.random((h, w, 3))
    L, d = pm.lkm_laplacian(image, epsilon=1e-6, radius=r)
    Lr, dr = olkm.lkm_laplacian(image, epsilon=1e-6, radius=r)
    assert np.abs(d - dr).max() <= 1e-10 * np.abs(dr).max()
    for _ in range(2):
        p = rng.standard_normal(h * w)
        ref = Lr(p)
        assert np.abs(L(p) - ref).max() <= 1e-10 * np.abs(ref).max()
    if r <= 2:
        C = ocf.cf_laplacian(image, epsilon=1e-6, radius=r)
        assert np.abs(L(p) - C @ p).max() <= 1e-8 * np.abs(C @ p).max()
    with pytest.raises(ValueError, match="smaller than one"):
        pm.lkm_laplacian(image[: 2 * r], radius=r)


def test_lemur_known_answer():
    import pymatting_amd as pm
    d = load_golden("lemur_0p125")
    image, trimap = d["image"], d["trimap"]
    L_lkm, diag = pm.lkm_laplacian(image, epsilon=1e-7, radius=2)
    A_cf, b, c = pm.make_linear_system(pm.cf_laplacian(image, epsilon=1e-7, radius=2), trimap,
                                       return_c=True)
    inv = 1.0 / (diag + c)
    n_lkm, n_cf = pm.CounterCallback(), pm.CounterCallback()
    x_lkm = pm.cg(lambda x: L_lkm(x) + c * x, b, M=lambda r: inv * r, callback=n_lkm)
    x_cf = pm.cg(A_cf, b, M=pm.jacobi(A_cf), callback=n_cf)
    assert np.linalg.norm(x_lkm - x_cf) < 1e-5
    assert abs(n_lkm.n - n_cf.n) <= 2


def test_estimate_alpha_lkm_matches_oracle():
    import pymatting_amd as pm
    import torch
    d = load_golden("lemur_0p125")
    image, trimap = d["image"], d["trimap"]
    kw = dict(laplacian_kwargs={"radius": 3, "epsilon": 1e-6}, cg_kwargs={"rtol": 1e-10})
    alpha = pm.estimate_alpha_lkm(image, trimap, **kw)
    alpha_ref = olkm.estimate_alpha_lkm(image, trimap, **kw)
    assert np.abs(alpha - alpha_ref).max() <= 1e-6
    at = pm.estimate_alpha_lkm(torch.from_numpy(image).cuda(), torch.from_numpy(trimap).cuda(),
                               **kw)
    assert at.is_cuda and float((at.cpu() - torch.from_numpy(alpha)).abs().max()) <= 1e-12


def test_cutout(tmp_path):
    """tests/test_simple_api.py:6-27 on PNG files written from the fixture."""
    import pymatting_amd as pm
    t = load_golden("test32")
    img_path, tri_path = str(tmp_path / "test.png"), str(tmp_path / "trimap.png")
    pm.save_image(img_path, np.round(t["image"] * 255).astype(np.uint8))
    pm.save_image(tri_path, np.round(t["trimap"] * 255).astype(np.uint8))
    out = str(tmp_path / "out" / "cutout.png")
    pm.cutout(img_path, tri_path, out)
    cut = pm.load_image(out)
    foreground, alpha = cut[:, :, :3], cut[:, :, 3]
    assert np.linalg.norm(alpha - t["true_alpha"]) < 3.1
    assert np.linalg.norm(alpha[:, :, None] * np.abs(foreground - t["true_foreground"])) < 4
    with pytest.raises(ValueError, match="same size"):
        small = str(tmp_path / "small.png")
        pm.save_image(small, np.zeros((8, 8), np.uint8))
        pm.cutout(img_path, small, out)
