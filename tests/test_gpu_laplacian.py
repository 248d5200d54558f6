"""GPU parity of the closed-form Laplacian against the oracle.

Entries of cf_laplacian must be bit-identical (the library gathers each entry
in the reference's accumulation order with no FMA contraction); the
matrix-free operator must agree with the oracle's CSR product to 1e-12
relative.
"""
import numpy as np
import pytest
import scipy.sparse

from conftest import load_golden
import oracle.cf as ocf

pytestmark = pytest.mark.gpu


def _images():
    rng = np.random.default_rng(21)
    yield "rand_5x7", rng.random((5, 7, 3)), None
    yield "rand_33x17", rng.random((33, 17, 3)), None
    yield "flat", np.full((6, 9, 3), 0.3), None
    d = load_golden("lemur_0p1")
    yield "lemur", d["image"], None
    yield "lemur_known", d["image"], ocf.trimap_split(d["trimap"])[2]
    yield "tiny_2x2", rng.random((2, 2, 3)), None
    yield "row_1x9", rng.random((1, 9, 3)), None


@pytest.mark.parametrize("case", list(range(7)))
def test_cf_laplacian_bit_exact(case):
    import pymatting_amd as pm
    name, image, known = list(_images())[case]
    print("case", name, image.shape, flush=True)
    L = pm.cf_laplacian(image, is_known=known)
    R = ocf.cf_laplacian(image, is_known=known)
    assert np.array_equal(L.indptr, R.indptr)
    assert np.array_equal(L.indices, R.indices)
    assert np.array_equal(L.data, R.data), np.abs(L.data - R.data).max()


def test_make_linear_system_canonical_and_exact():
    import pymatting_amd as pm
    d = load_golden("lemur_0p125")
    A, b = pm.make_linear_system(pm.cf_laplacian(d["image"]), d["trimap"])
    Ar, br = ocf.make_linear_system(ocf.cf_laplacian(d["image"]), d["trimap"])
    assert A.has_canonical_format
    assert np.array_equal(A.indptr, Ar.indptr)
    assert np.array_equal(A.indices, Ar.indices)
    assert np.array_equal(A.data, Ar.data)
    assert np.array_equal(b, br)


@pytest.mark.parametrize("mode", ["laplacian", "lambda", "masked"])
def test_matrix_free_apply(mode):
    import torch
    import pymatting_amd as pm
    from pymatting_amd.util import device_classes
    d = load_golden("lemur_0p125")
    image, trimap = d["image"], d["trimap"]
    n = trimap.size
    Lr = ocf.cf_laplacian(image)
    fg, bg, known, unknown = ocf.trimap_split(trimap)
    if mode == "laplacian":
        Ar, cls = Lr, None
    elif mode == "lambda":
        Ar = ocf.make_linear_system(Lr, trimap)[0]
    else:
        Ad = Lr.toarray()
        Ad[known, :] = 0
        Ad[:, known] = 0
        Ad[known, known] = 1.0
        Ar = Ad
    cls = None if mode == "laplacian" else device_classes(trimap)[0]
    op = pm.CfOperator(image, mode=mode, classes=cls, lambda_value=100.0 if mode == "lambda" else 0)
    rng = np.random.default_rng(0)
    absA = abs(scipy.sparse.csr_matrix(Ar))
    for _ in range(3):
        x = rng.standard_normal(n)
        y = op.dot(x)
        yr = Ar.dot(x)
        # The factored stencil rounds differently from the CSR product; the
        # bound scales with |A| |x| (the window inverses are bit-identical).
        scale = (absA @ np.abs(x)).max()
        assert np.abs(y - yr).max() <= 1e-10 * scale
        y32 = op.apply(torch.from_numpy(x.astype(np.float32)).cuda(), precision="f32").cpu().numpy()
        assert np.abs(y32 - yr).max() <= 1e-3 * scale
    dr = Ar.diagonal() if mode != "masked" else np.diag(Ar)
    assert np.array_equal(op.diagonal(), dr)


def test_masked_rhs_exact():
    import pymatting_amd as pm
    from pymatting_amd.util import device_classes
    d = load_golden("lemur_0p1")
    image, trimap = d["image"], d["trimap"]
    fg, bg, known, unknown = ocf.trimap_split(trimap)
    L = ocf.cf_laplacian(image, is_known=known)
    R = L[unknown, :][:, known]
    ref = -R.dot(fg[known])
    cls, t = device_classes(trimap)
    op = pm.CfOperator(image, mode="masked", classes=cls)
    b = op.rhs().cpu().numpy()
    assert np.array_equal(b[unknown], ref)
    assert np.all(b[known] == 0)


def _slot_array(A, h, w):
    """[25][n] stencil array of a CSR matrix on an (h, w) grid (oracle side).
    The reference's CSR keeps 25 entries per row, out-of-image slots as
    explicit zeros in column 0; zeros are dropped here."""
    A = scipy.sparse.coo_matrix(A)
    keep = A.data != 0
    A = scipy.sparse.coo_matrix((A.data[keep], (A.row[keep], A.col[keep])), shape=A.shape)
    n = h * w
    dy = A.col // w - A.row // w
    dx = A.col % w - A.row % w
    assert np.all(np.abs(dy) <= 2) and np.all(np.abs(dx) <= 2)
    coef = np.zeros((25, n))
    coef[(dy + 2) * 5 + dx + 2, A.row] = A.data
    return coef


@pytest.mark.parametrize("case", list(range(7)))
def test_stored_stencil_bit_exact(case):
    """The stored 13-slot form assembled from the window data equals the
    oracle's CSR entries of those slots bit for bit in every operator mode: the Laplacian,
    L + lambda diag(known) of make_linear_system, and the masked system
    [L_UU 0; 0 I] of estimate_alpha_cf (odd sizes cross the tile edges)."""
    import torch
    from pymatting_amd.laplacian import CfOperator, DiaOperator
    from pymatting_amd.util import device_classes
    name, image, _ = list(_images())[case]
    h, w = image.shape[:2]
    rng = np.random.default_rng(case)
    trimap = rng.choice([0.0, 0.5, 1.0], size=(h, w), p=[0.3, 0.4, 0.3])
    trimap.flat[0], trimap.flat[-1] = 0.0, 1.0  # both known classes present
    print("case", name, image.shape, flush=True)
    L = ocf.cf_laplacian(image)
    A, _ = ocf.make_linear_system(L, trimap)
    known = ocf.trimap_split(trimap)[2]
    M = scipy.sparse.lil_matrix(L)
    M[known, :] = 0
    M[:, known] = 0
    M = scipy.sparse.csr_matrix(M) + scipy.sparse.diags(known.astype(np.float64))
    classes, tdev = device_classes(torch.from_numpy(trimap).cuda())
    for mode, ref in (("laplacian", L), ("lambda", A), ("masked", M)):
        kw = {} if mode == "laplacian" else dict(classes=classes, trimap=tdev)
        cf = CfOperator(image, mode=mode, lambda_value=100.0 if mode == "lambda" else 0.0, **kw)
        coef = DiaOperator.from_cf(cf).coef.cpu().numpy()
        expect = _slot_array(ref, h, w)
        # Stored slots 12..24 ((dy, dx) >= (0, 0)) are bit-exact; the other
        # twelve mirror them: slot s at i is the reference's L_ij for
        # j = i + (dy, dx), i.e. slot 24 - s at j (the reference's own L_ji
        # rounds differently).
        assert np.array_equal(coef[12:], expect[12:]), \
            (mode, np.abs(coef[12:] - expect[12:]).max())
        mirror = np.zeros((12, h * w))
        for s in range(12):
            dy, dx = s // 5 - 2, s % 5 - 2
            for y in range(max(0, -dy), min(h, h - dy)):
                for x in range(max(0, -dx), min(w, w - dx)):
                    mirror[s, y * w + x] = expect[24 - s, (y + dy) * w + x + dx]
        assert np.array_equal(coef[:12], mirror), mode
