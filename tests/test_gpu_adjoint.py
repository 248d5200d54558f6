"""The hand-written adjoint of the alpha solve (pymatting_amd.autograd)
against the oracle's autograd gradient of a dense restatement and against
central finite differences of the GPU forward solve."""
import numpy as np
import pytest

from oracle.adjoint import alpha_loss_and_gradient

pytestmark = pytest.mark.gpu


def _trimap(h, w):
    t = np.full((h, w), 0.5)
    t[:, :2] = 0.0
    t[:, -2:] = 1.0
    return t


def _grad(image, trimap, target, **opts):
    import torch
    import pymatting_amd as pm
    I = torch.tensor(image, device="cuda", requires_grad=True)
    alpha = pm.solve_alpha(I, torch.tensor(trimap, device="cuda"), **opts)
    loss = ((alpha - torch.tensor(target, device="cuda", dtype=alpha.dtype)) ** 2).mean()
    loss.backward()
    return float(loss.detach()), alpha.detach().cpu().numpy(), I.grad.cpu().numpy()


def test_adjoint_matches_reference_gradient():
    rng = np.random.default_rng(11)
    h, w = 12, 14
    image = rng.random((h, w, 3))
    trimap = _trimap(h, w)
    target = rng.random((h, w))
    loss_r, alpha_r, grad_r = alpha_loss_and_gradient(image, trimap, target)
    # direct_solve_size >= h*w: the preconditioner is the exact inverse.
    loss, alpha, grad = _grad(image, trimap, target, precision="f64", rtol=1e-12,
                              direct_solve_size=256)
    assert np.abs(alpha - alpha_r).max() < 1e-9
    assert abs(loss - loss_r) < 1e-10
    assert np.abs(grad - grad_r).max() <= 1e-7 * np.abs(grad_r).max()


def test_adjoint_finite_differences():
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    h, w = 40, 48
    image, trimap, a_true = synthetic.make_image(h, w, 5004, strands=4, erode=3)
    target = np.clip(a_true + np.random.default_rng(1).normal(0, 0.05, a_true.shape), 0, 1)
    opts = dict(precision="f64", rtol=1e-13)
    loss, _, grad = _grad(image, trimap, target, **opts)
    tt = torch.tensor(trimap, device="cuda")
    tg = torch.tensor(target, device="cuda")

    def loss_at(img):
        with torch.no_grad():
            a = pm.solve_alpha(torch.tensor(img, device="cuda"), tt, **opts)
            return float(((a - tg) ** 2).mean())

    unknown = np.argwhere(trimap == 0.5)
    rng = np.random.default_rng(2)
    step = 1e-3
    for k in rng.choice(len(unknown), 20, replace=False):
        y, x = unknown[k]
        c = int(rng.integers(0, 3))
        ip, im = image.copy(), image.copy()
        ip[y, x, c] += step
        im[y, x, c] -= step
        fd = (loss_at(ip) - loss_at(im)) / (2 * step)
        assert abs(fd - grad[y, x, c]) <= 2e-3 * np.abs(grad).max() + 2e-2 * abs(fd), (y, x, c)


def test_batched_float32_mixed_gradient():
    import torch
    from pymatting_amd import synthetic
    imgs, tris, tgts = [], [], []
    for i in range(2):
        image, trimap, a = synthetic.make_image(64, 80, 5100 + i, strands=6, erode=4)
        imgs.append(image)
        tris.append(trimap)
        tgts.append(np.clip(a + np.random.default_rng(i).normal(0, 0.05, a.shape), 0, 1))
    image = np.stack(imgs)
    trimap = np.stack(tris)
    target = np.stack(tgts)
    loss32, alpha32, grad32 = _grad(image.astype(np.float32), trimap, target)
    loss64, alpha64, grad64 = _grad(image, trimap, target, precision="f64", rtol=1e-12,
                                    direct_solve_size=64)
    assert grad32.dtype == np.float32 and grad32.shape == image.shape
    assert np.abs(alpha32 - alpha64).max() < 1e-4
    assert np.linalg.norm(grad32 - grad64) <= 1e-3 * np.linalg.norm(grad64)


def test_batched_and_per_image_gradients_agree():
    """The batched forward and adjoint PCG (one BatchSolver for the batch)
    give the gradient of the image-by-image path; an image whose loss
    gradient is zero (alpha equal to its target) gets a zero image gradient."""
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    imgs, tris = [], []
    for i in range(3):
        image, trimap, _ = synthetic.make_image(48, 72, 5200 + i, strands=4, erode=4)
        imgs.append(image)
        tris.append(trimap)
    image = torch.tensor(np.stack(imgs), device="cuda")
    trimap = torch.tensor(np.stack(tris), device="cuda")
    opts = dict(precision="f64", rtol=1e-12)
    with torch.no_grad():
        alpha0 = pm.solve_alpha(image, trimap, batched=False, **opts)
    target = alpha0.clone()
    target[:2] += 0.05 * torch.randn_like(target[:2])  # image 2 keeps its own alpha

    def grad(batched):
        I = image.clone().requires_grad_(True)
        alpha = pm.solve_alpha(I, trimap, batched=batched, **opts)
        ((alpha - target) ** 2).mean().backward()
        return alpha.detach(), I.grad

    a_b, g_b = grad(True)
    a_s, g_s = grad(False)
    print("alpha", float((a_b - a_s).abs().max()), "grad", float((g_b - g_s).norm() / g_s.norm()),
          flush=True)
    assert float((a_b - a_s).abs().max()) < 1e-8
    assert float((g_b - g_s).norm()) <= 1e-7 * float(g_s.norm())
    # Image 2's target is its per-image alpha: the batched alpha differs from
    # it only by rounding, so its gradient is that small.
    assert float(g_b[2].abs().max()) <= 1e-9 * float(g_b.abs().max())
