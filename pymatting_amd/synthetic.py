"""Deterministic synthetic matting workloads (host-side test/benchmark data).

image = a*F + (1-a)*B with
  - F, B: per channel 0.5 + a sum of 4 random low-frequency sinusoids (at most
    3 cycles per image, amplitudes summing to at most 0.45) plus Gaussian
    noise sigma 0.02, clipped to [0, 1];
  - a: a filled ellipse covering ~30 % of the image united with thin
    random-walk strands (width 1-2 px, length ~0.3*max(h, w)), blurred with a
    Gaussian of sigma 3 px;
  - trimap: 1 where (a > 0.99) eroded by `erode` px, 0 where (a < 0.01)
    eroded by `erode` px, 0.5 elsewhere. An "unknown band of B px" uses
    erode = B // 2 on both sides of the blurred edge.
Seeds: seed = 1000 * config + image index.
"""
import numpy as np
import scipy.ndimage


def _layer(rng, h, w):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    out = np.empty((h, w, 3))
    for c in range(3):
        v = np.full((h, w), 0.5)
        amps = rng.uniform(0.0, 1.0, 4)
        amps *= 0.45 / amps.sum()
        for k in range(4):
            fx, fy = rng.uniform(-3.0, 3.0, 2)
            phase = rng.uniform(0, 2 * np.pi)
            v += amps[k] * np.sin(2 * np.pi * (fx * x / w + fy * y / h) + phase)
        out[:, :, c] = v
    out += rng.normal(0.0, 0.02, out.shape)
    return np.clip(out, 0.0, 1.0)


def _alpha(rng, h, w, strands):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    area = 0.3 * h * w
    ratio = rng.uniform(0.6, 1.0)
    a_ax = np.sqrt(area / (np.pi * ratio))
    b_ax = a_ax * ratio
    cy = rng.uniform(0.35, 0.65) * h
    cx = rng.uniform(0.35, 0.65) * w
    th = rng.uniform(0, np.pi)
    u = (x - cx) * np.cos(th) + (y - cy) * np.sin(th)
    v = -(x - cx) * np.sin(th) + (y - cy) * np.cos(th)
    mask = (u / a_ax) ** 2 + (v / b_ax) ** 2 <= 1.0
    length = int(0.3 * max(h, w))
    for _ in range(strands):
        # Strands start on the ellipse boundary and wander outwards.
        t = rng.uniform(0, 2 * np.pi)
        px = cx + a_ax * np.cos(t) * np.cos(th) - b_ax * np.sin(t) * np.sin(th)
        py = cy + a_ax * np.cos(t) * np.sin(th) + b_ax * np.sin(t) * np.cos(th)
        width = int(rng.integers(1, 3))
        dx, dy = np.cos(t), np.sin(t)
        for _ in range(length):
            dx += rng.normal(0, 0.3)
            dy += rng.normal(0, 0.3)
            nrm = np.hypot(dx, dy) + 1e-12
            dx, dy = dx / nrm, dy / nrm
            px, py = px + dx, py + dy
            ix, iy = int(round(px)), int(round(py))
            if not (0 <= ix < w and 0 <= iy < h):
                break
            mask[iy:iy + width, ix:ix + width] = True
    return np.clip(scipy.ndimage.gaussian_filter(mask.astype(np.float64), 3.0), 0.0, 1.0)


def _trimap(alpha, erode):
    fg = alpha > 0.99
    bg = alpha < 0.01
    if erode > 0:
        fg = scipy.ndimage.distance_transform_edt(fg) > erode
        bg = scipy.ndimage.distance_transform_edt(bg) > erode
    t = np.full(alpha.shape, 0.5)
    t[fg] = 1.0
    t[bg] = 0.0
    return t


def make_image(h, w, seed, strands=20, erode=8):
    """One (image, trimap, true_alpha) triple, float64."""
    rng = np.random.default_rng(seed)
    F = _layer(rng, h, w)
    B = _layer(rng, h, w)
    a = _alpha(rng, h, w, strands)
    image = a[:, :, None] * F + (1.0 - a[:, :, None]) * B
    return image, _trimap(a, erode), a


# The five benchmark configurations.
def config1():
    """320x240, 20 strands, 8 px erosion."""
    return make_image(240, 320, 1000, strands=20, erode=8)


def config2():
    """1920x1080, 40 strands, unknown band 20 px."""
    return make_image(1080, 1920, 2000, strands=40, erode=10)


def config3(count=64):
    """64 images, h, w uniform in 256..1024 step 16, bands 5-40 px."""
    rng = np.random.default_rng(3000)
    out = []
    for i in range(count):
        h = int(rng.integers(0, 49)) * 16 + 256
        w = int(rng.integers(0, 49)) * 16 + 256
        band = int(rng.integers(5, 41))
        out.append(make_image(h, w, 3001 + i, strands=20, erode=band // 2))
    return out


def config4():
    """3840x2160, 20 strands, unknown band 40 px."""
    return make_image(2160, 3840, 4000, strands=20, erode=20)


def config5(count=16):
    """16 images at 512x512 with a noisy target alpha (sigma 0.05)."""
    rng = np.random.default_rng(5000)
    out = []
    for i in range(count):
        image, trimap, a = make_image(512, 512, 5001 + i, strands=20, erode=8)
        out.append((image, trimap, a + rng.normal(0.0, 0.05, a.shape)))
    return out
