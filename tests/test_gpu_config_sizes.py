"""Parity at the configurations' own sizes.

- 1920x1080: cf_laplacian's entries bit-identical to the oracle's; the
  float64 GPU lambda-system solve (V(1,1) V-cycle PCG, rtol 1e-7) within
  1e-6 of the oracle's alpha, in the oracle's iteration count (+-1), and the
  mixed form within the metric's 1e-4 of both.
- 3840x2160: estimate_foreground_ml in float32 within 1e-4 of float64 on a
  solved alpha.
- 16 images of 512x512 (float32 tensors): the hand-written adjoint of the
  batched mixed solve against central finite differences (step 1e-3,
  float64 solves) on 200 random pixel channels, and against the float64
  adjoint gradient.

The oracle's 1080p solve takes about 10 s on one core.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_1080p_laplacian_and_fp64_alpha_match_the_oracle():
    import torch
    import pymatting_amd as pm
    import oracle.cf as ocf
    from oracle.solver import CounterCallback, cg as ocg, vcycle as ovcycle
    from pymatting_amd import synthetic
    image, trimap, _ = synthetic.config2()
    h, w = trimap.shape
    print("1080p oracle cf_laplacian", flush=True)
    Lr = ocf.cf_laplacian(image)
    print("1080p GPU cf_laplacian", flush=True)
    L = pm.cf_laplacian(image)
    assert np.array_equal(L.indptr, Lr.indptr) and np.array_equal(L.indices, Lr.indices)
    assert np.array_equal(L.data, Lr.data)
    del L
    print("1080p oracle solve", flush=True)
    A, b = ocf.make_linear_system(Lr, trimap)
    del Lr
    counter = CounterCallback()
    x_ref = ocg(A, b, rtol=1e-7, M=ovcycle(A, (h, w)), callback=counter)
    alpha_ref = np.clip(x_ref, 0, 1).reshape(h, w)
    del A
    print("1080p GPU solves (oracle: %d iterations)" % counter.n, flush=True)
    img_d, tri_d = torch.from_numpy(image).cuda(), torch.from_numpy(trimap).cuda()
    f64 = pm.LambdaSolver(img_d, tri_d, precision="f64", level0="stencil")
    a64 = f64.solve(rtol=1e-7).cpu().numpy()
    assert np.abs(a64 - alpha_ref).max() <= 1e-6, np.abs(a64 - alpha_ref).max()
    assert abs(f64.stats["iterations"] - counter.n) <= 1
    mixed = pm.LambdaSolver(img_d, tri_d, precision="mixed", level0="stencil")
    am = mixed.solve(rtol=1e-7, mixed="reliable").cpu().numpy()
    assert np.abs(am - a64).max() <= 1e-4 and np.abs(am - alpha_ref).max() <= 1e-4
    # The drop-in chain on the same image: cf_laplacian -> make_linear_system
    # -> cg(M=vcycle) on the device.
    A, b = pm.make_linear_system(pm.cf_laplacian(image), trimap)
    x = pm.cg(A, b, rtol=1e-7, M=pm.vcycle(A, (h, w)))
    assert np.abs(np.clip(x, 0, 1).reshape(h, w) - alpha_ref).max() <= 1e-6


def test_4k_float32_foreground_within_1e4_of_float64():
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    image_h, trimap_h, _ = synthetic.config4()
    image = torch.from_numpy(image_h).cuda()
    trimap = torch.from_numpy(trimap_h).cuda()
    print("4K alpha", flush=True)
    s = pm.LambdaSolver(image, trimap, precision="mixed", level0="stencil", num_pre_iter=2,
                        num_post_iter=2, max_levels=5, coarse_sweeps=2)
    alpha = s.solve(rtol=1e-7, mixed="reliable")
    print("4K foreground", flush=True)
    F32, B32 = pm.estimate_foreground_ml(image.float(), alpha.float(), return_background=True)
    F64, B64 = pm.estimate_foreground_ml(image, alpha, return_background=True, precision="f64")
    assert F32.dtype == torch.float32 and F32.shape == (2160, 3840, 3)
    assert float((F32.double() - F64).abs().max()) <= 1e-4
    assert float((B32.double() - B64).abs().max()) <= 1e-4


def test_c5_batch_gradient_finite_differences_and_fp64():
    import torch
    import pymatting_amd as pm
    from pymatting_amd import synthetic
    data = synthetic.config5()
    images = np.stack([d[0] for d in data])
    trimaps = np.stack([d[1] for d in data])
    targets = np.stack([d[2] for d in data])
    B, h, w, _ = images.shape
    assert (B, h, w) == (16, 512, 512)
    tri = torch.from_numpy(trimaps).cuda()
    tgt = torch.from_numpy(targets).cuda()

    def grad(img, **opts):
        x = img.clone().requires_grad_(True)
        alpha = pm.solve_alpha(x, tri, **opts)
        ((alpha.double() - tgt) ** 2).mean().backward()
        return x.grad.double()

    print("C5 batched mixed forward/backward", flush=True)
    g32 = grad(torch.from_numpy(images.astype(np.float32)).cuda())
    assert g32.shape == (B, h, w, 3)
    img64 = torch.from_numpy(images).cuda()
    print("C5 float64 adjoint", flush=True)
    g64 = grad(img64, precision="f64", rtol=1e-12)
    rel = float((g32 - g64).norm() / g64.norm())
    assert rel <= 1e-3, rel
    print("C5 finite differences", flush=True)
    opts = dict(precision="f64", rtol=1e-13)
    n_px = B * h * w
    rng = np.random.default_rng(5)
    step = 1e-3
    gmax = float(g64.abs().max())
    worst = 0.0
    with torch.no_grad():
        def loss_of(i, img_i):
            a = pm.solve_alpha(img_i, tri[i], **opts)
            return float(((a - tgt[i]) ** 2).sum()) / n_px
        for _ in range(200):
            i = int(rng.integers(0, B))
            unk = np.argwhere(trimaps[i] == 0.5)
            y, x = unk[int(rng.integers(0, len(unk)))]
            c = int(rng.integers(0, 3))
            ip, im = img64[i].clone(), img64[i].clone()
            ip[y, x, c] += step
            im[y, x, c] -= step
            fd = (loss_of(i, ip) - loss_of(i, im)) / (2 * step)
            worst = max(worst, abs(fd - float(g64[i, y, x, c])) / gmax)
    assert worst <= 1e-3, worst
