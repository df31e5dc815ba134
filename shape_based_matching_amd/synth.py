"""Seeded synthetic scenes and template banks for the benchmark workloads.

Workloads (all uint8 grayscale, numpy PCG64 seeded; the seed is part of every result record):

* ``scene_plain``   640x480: background 128 + N(0, 10^2), 20 filled ellipses / polygons (5-12 vertices)
                    of intensity U[30, 225]; the template is a 160x160 crop around one polygon.
* ``scene_copies``  1024x1024 with 8 rotated (U[0,360)), scaled (U[0.9,1.1]) copies of the template
                    polygon, Gaussian blur sigma ~ U[0, 2] and 2% salt-and-pepper noise.
* ``synthetic_bank``  classes x rotations x scales of random polygons with features placed on the
                    transformed outline (labels from the edge normal); level-1 count ceil(n/2).
"""
import math

import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


# ------------------------------------------------------------------------------------------------
# rasterisation
# ------------------------------------------------------------------------------------------------
def random_polygon(g, nv=None, radius=1.0):
    """Star-shaped polygon with nv vertices (5..12) around the origin, radii in [0.55, 1] * radius."""
    nv = int(g.integers(5, 13)) if nv is None else nv
    ang = np.sort(g.uniform(0, 2 * np.pi, nv))
    rad = g.uniform(0.55, 1.0, nv) * radius
    return np.stack([rad * np.cos(ang), rad * np.sin(ang)], axis=1)


def transform(poly, angle_deg, scale, cx, cy):
    a = math.radians(angle_deg)
    c, s = math.cos(a), math.sin(a)
    R = np.array([[c, -s], [s, c]]) * scale
    return poly @ R.T + np.array([cx, cy])


def fill_polygon(img, poly, value):
    """Even-odd fill of pixel centres inside poly (x, y vertex array)."""
    h, w = img.shape
    x0 = max(int(np.floor(poly[:, 0].min())), 0)
    x1 = min(int(np.ceil(poly[:, 0].max())) + 1, w)
    y0 = max(int(np.floor(poly[:, 1].min())), 0)
    y1 = min(int(np.ceil(poly[:, 1].max())) + 1, h)
    if x0 >= x1 or y0 >= y1:
        return
    ys, xs = np.mgrid[y0:y1, x0:x1]
    px, py = xs + 0.5, ys + 0.5
    inside = np.zeros(px.shape, bool)
    n = len(poly)
    for i in range(n):
        xa, ya = poly[i]
        xb, yb = poly[(i + 1) % n]
        if ya == yb:
            continue
        cond = (ya > py) != (yb > py)
        xint = xa + (py - ya) * (xb - xa) / (yb - ya)
        inside ^= cond & (px < xint)
    img[y0:y1, x0:x1][inside] = value


def fill_ellipse(img, cx, cy, rx, ry, angle_deg, value):
    h, w = img.shape
    r = max(rx, ry)
    x0, x1 = max(int(cx - r) - 1, 0), min(int(cx + r) + 2, w)
    y0, y1 = max(int(cy - r) - 1, 0), min(int(cy + r) + 2, h)
    if x0 >= x1 or y0 >= y1:
        return
    ys, xs = np.mgrid[y0:y1, x0:x1]
    a = math.radians(angle_deg)
    dx, dy = xs + 0.5 - cx, ys + 0.5 - cy
    u = dx * math.cos(a) + dy * math.sin(a)
    v = -dx * math.sin(a) + dy * math.cos(a)
    img[y0:y1, x0:x1][(u / rx) ** 2 + (v / ry) ** 2 <= 1.0] = value


def noisy_background(g, rows, cols, mean=128.0, sigma=10.0):
    return np.clip(np.rint(mean + g.normal(0, sigma, (rows, cols))), 0, 255).astype(np.uint8)


def gaussian_blur(img, sigma):
    if sigma <= 0.05:
        return img.copy()
    r = max(1, int(math.ceil(3 * sigma)))
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
    k /= k.sum()
    f = img.astype(np.float64)
    p = np.pad(f, ((0, 0), (r, r)), mode="edge")
    f = sum(k[i] * p[:, i:i + img.shape[1]] for i in range(2 * r + 1))
    p = np.pad(f, ((r, r), (0, 0)), mode="edge")
    f = sum(k[i] * p[i:i + img.shape[0], :] for i in range(2 * r + 1))
    return np.clip(np.rint(f), 0, 255).astype(np.uint8)


def salt_pepper(g, img, frac):
    out = img.copy()
    m = g.random(img.shape) < frac
    out[m] = np.where(g.random(int(m.sum())) < 0.5, 0, 255).astype(np.uint8)
    return out


# ------------------------------------------------------------------------------------------------
# the plain-method scene and its template
# ------------------------------------------------------------------------------------------------
def scene_plain(seed=0, rows=480, cols=640, n_shapes=20):
    """640x480 scene; returns (scene, template_image 160x160, template_polygon (unit), intensity)."""
    g = rng(seed)
    img = noisy_background(g, rows, cols)
    templ_poly = random_polygon(g, nv=int(g.integers(5, 13)), radius=1.0)
    templ_val = int(g.integers(30, 60))  # dark polygon: strong contrast with the background
    for i in range(n_shapes):
        cx, cy = g.uniform(40, cols - 40), g.uniform(40, rows - 40)
        val = int(g.integers(30, 226))
        if i == 0:
            fill_polygon(img, transform(templ_poly, 0, 60, cx, cy), templ_val)
        elif g.random() < 0.5:
            fill_ellipse(img, cx, cy, g.uniform(8, 40), g.uniform(8, 40), g.uniform(0, 180), val)
        else:
            fill_polygon(img, transform(random_polygon(g), g.uniform(0, 360), g.uniform(10, 45), cx, cy), val)
    templ = template_image(seed, templ_poly, templ_val)
    return img, templ, templ_poly, templ_val


def template_image(seed, poly, val, size=160, radius=60):
    g = rng(seed + 7919)
    t = noisy_background(g, size, size)
    fill_polygon(t, transform(poly, 0, radius, size / 2, size / 2), val)
    return t


def scene_copies(seed=1, size=1024, n_copies=8, templ_seed=0, radius=60):
    """1024x1024 scene with rotated/scaled copies of the template polygon; returns (scene, placements)."""
    _, templ, poly, val = scene_plain(templ_seed)
    g = rng(seed)
    img = noisy_background(g, size, size)
    placed = []
    tries = 0
    while len(placed) < n_copies and tries < 10000:
        tries += 1
        sc = g.uniform(0.9, 1.1)
        r = radius * sc + 6
        cx, cy = g.uniform(r + 2, size - r - 2), g.uniform(r + 2, size - r - 2)
        if any((cx - p[0]) ** 2 + (cy - p[1]) ** 2 < (r + p[3]) ** 2 for p in placed):
            continue
        ang = g.uniform(0, 360)
        fill_polygon(img, transform(poly, ang, radius * sc, cx, cy), val)
        placed.append((cx, cy, ang, r, sc))
    img = gaussian_blur(img, g.uniform(0, 2))
    img = salt_pepper(g, img, 0.02)
    return img, templ, placed


# ------------------------------------------------------------------------------------------------
# synthetic banks (features on transformed outlines)
# ------------------------------------------------------------------------------------------------
def outline_features(poly_xy, n, bins=8):
    """n points spread uniformly along the closed outline, with labels from the edge normal (bins
    orientations over 180 degrees: 8, or 16 for the 16-orientation variant)."""
    p = np.asarray(poly_xy, np.float64)
    q = np.roll(p, -1, axis=0)
    seg = np.hypot(*(q - p).T)
    cum = np.concatenate([[0.0], np.cumsum(seg)])
    t = (np.arange(n) + 0.5) * (cum[-1] / n)
    k = np.clip(np.searchsorted(cum, t, side="right") - 1, 0, len(p) - 1)
    f = (t - cum[k]) / np.maximum(seg[k], 1e-9)
    pts = p[k] + (q[k] - p[k]) * f[:, None]
    d = q[k] - p[k]
    normal = np.degrees(np.arctan2(d[:, 0], -d[:, 1])) % 360.0  # perpendicular to the edge
    label = (np.floor(normal * (2 * bins) / 360 + 0.5).astype(np.int64)) & (bins - 1)
    return np.rint(pts).astype(np.int64), label


def bank_template(poly_unit, angle, scale, radius, n0, bins=8):
    """Two-level pyramid (dicts) for one transformed polygon with n0 level-0 features."""
    outline = transform(poly_unit, angle, radius * scale, 0.0, 0.0)
    pts0, lab0 = outline_features(outline, n0, bins)
    n1 = (n0 + 1) // 2
    pts1, lab1 = outline_features(outline / 2.0, n1, bins)
    mn0 = pts0.min(axis=0)
    mn0 -= mn0 % 2  # even top-left, as cropTemplates (line2Dup.cpp:126-130)
    mn1 = mn0 // 2
    f0 = np.column_stack([pts0 - mn0, lab0]).astype(np.int32)
    f1 = np.column_stack([pts1 - mn1, lab1]).astype(np.int32)
    f1[:, :2] = np.maximum(f1[:, :2], 0)
    w0 = int(f0[:, 0].max()) if len(f0) else 0
    h0 = int(f0[:, 1].max()) if len(f0) else 0
    lv0 = dict(width=w0, height=h0, tl_x=int(mn0[0]), tl_y=int(mn0[1]), features=f0)
    lv1 = dict(width=w0 >> 1, height=h0 >> 1, tl_x=int(mn1[0]), tl_y=int(mn1[1]), features=f1)
    return [lv0, lv1]


def synthetic_bank(seed=2, n_classes=20, angles=range(0, 360), scales=(0.8, 0.9, 1.0, 1.1, 1.2), radius=60,
                   nmin=32, nmax=1024, big_frac=0.01, big_n=8191):
    """Yields (class_id, pyramid, polygon_unit) for classes x scales x angles templates."""
    g = rng(seed)
    polys = [random_polygon(g) for _ in range(n_classes)]
    for ci in range(n_classes):
        cid = f"poly{ci:02d}"
        for sc in scales:
            for ang in angles:
                n0 = big_n if g.random() < big_frac else int(g.integers(nmin, nmax + 1))
                yield cid, bank_template(polys[ci], float(ang), float(sc), radius, n0), polys[ci]


def scene_with_bank_shapes(seed, size, polys, n_shapes=24, radius=60):
    g = rng(seed)
    img = noisy_background(g, size, size)
    for _ in range(n_shapes):
        poly = polys[int(g.integers(0, len(polys)))]
        sc = g.uniform(0.8, 1.2)
        cx, cy = g.uniform(radius * 1.3, size - radius * 1.3), g.uniform(radius * 1.3, size - radius * 1.3)
        fill_polygon(img, transform(poly, g.uniform(0, 360), radius * sc, cx, cy), int(g.integers(30, 80)))
    return img


# ------------------------------------------------------------------------------------------------
# mixed-size image batches
# ------------------------------------------------------------------------------------------------
BATCH_SIZES = ((480, 640), (960, 1280), (1936, 2592))  # 2592x1944 cropped to 1936 rows: every level divisible


def scene_sized(seed, rows, cols, poly, n_shapes=None, n_copies=3, radius=60):
    """Background + noise, 10-40 random shapes and a few rotated copies of `poly` (the bank's shape)."""
    g = rng(seed)
    img = noisy_background(g, rows, cols)
    n_shapes = int(g.integers(10, 41)) if n_shapes is None else n_shapes
    for _ in range(n_shapes):
        cx, cy = g.uniform(30, cols - 30), g.uniform(30, rows - 30)
        val = int(g.integers(30, 226))
        if g.random() < 0.5:
            fill_ellipse(img, cx, cy, g.uniform(8, 40), g.uniform(8, 40), g.uniform(0, 180), val)
        else:
            fill_polygon(img, transform(random_polygon(g), g.uniform(0, 360), g.uniform(10, 45), cx, cy), val)
    for _ in range(n_copies):
        cx, cy = g.uniform(radius + 4, cols - radius - 4), g.uniform(radius + 4, rows - radius - 4)
        fill_polygon(img, transform(poly, g.uniform(0, 360), radius, cx, cy), int(g.integers(30, 60)))
    return img


def mixed_batch(seed=4, n=64, sizes=BATCH_SIZES):
    """n scenes whose sizes are drawn uniformly from `sizes`, all containing the template polygon."""
    _, _, poly, _ = scene_plain(0)
    g = rng(seed)
    out = []
    for i in range(n):
        rows, cols = sizes[int(g.integers(0, len(sizes)))]
        out.append(scene_sized(seed * 1000 + i, rows, cols, poly))
    return out, poly
