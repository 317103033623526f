"""Seeded synthetic rectified stereo pairs with a piecewise-planar disparity field.

Generator definition (fixed once so every run uses identical bytes; ``sha256`` reports the digest):

* left  = uniform integer noise in [0, maxval] (numpy PCG64, ``seed``), then the reference's 3x3 box
  filter (``DenseStereo.BoxFilter3``, DenseStereo.cs:138-159: float(sum of the in-image window) times
  1/(cols*rows), round to nearest even, saturate).
* disparity = ``planes`` random planes d = a*x + b*y + c, each pixel taking the plane of its nearest
  Voronoi seed; rounded to integers and clipped to [0, dmax]. Plane offsets c are uniform in
  [0.1*dmax, 0.9*dmax]; slopes keep the plane within +-dmax/4 over the image.
* foreground (optional, ``fg_frac``): that fraction of the pixels (a disc per plane) is taken by extra
  fronto-parallel planes nearer than anything behind them, so they occlude.
* right = the left image forward-warped to x - d (the larger disparity wins a collision); a hole takes
  the nearer-in-row visible neighbour with the smaller disparity (the background); then Gaussian noise of
  ``sigma`` (same generator), round to nearest even, clamp to [0, maxval].
"""
import hashlib

import numpy as np


def box3(img, maxval):
    """Vectorised BoxFilter3 of a single-channel image (see module docstring)."""
    a = img.astype(np.int64)
    H, W = a.shape
    p = np.pad(a, 1)
    s = sum(p[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    rows = np.full(H, 3.0, np.float32)
    cols = np.full(W, 3.0, np.float32)
    if H > 1:
        rows[0] = rows[-1] = 2.0
    else:
        rows[0] = 1.0
    if W > 1:
        cols[0] = cols[-1] = 2.0
    else:
        cols[0] = 1.0
    inv = (np.float32(1.0) / (rows[:, None] * cols[None, :])).astype(np.float32)
    v = s.astype(np.float32) * inv
    return np.clip(np.rint(v), 0, maxval).astype(img.dtype)


def disparity_field(W, H, dmax, planes, rng, fg_frac=0.0):
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float32)
    seeds = np.stack([rng.uniform(0, W, planes), rng.uniform(0, H, planes)], 1).astype(np.float32)
    span = 0.25 * dmax
    a = rng.uniform(-span, span, planes) / max(W, 1)
    b = rng.uniform(-span, span, planes) / max(H, 1)
    c = rng.uniform(0.1 * dmax, 0.9 * dmax, planes)
    best = np.zeros((H, W), np.int32)
    bestd = np.full((H, W), np.inf, np.float32)
    for k in range(planes):
        dist = (xs - seeds[k, 0]) ** 2 + (ys - seeds[k, 1]) ** 2
        m = dist < bestd
        bestd[m] = dist[m]
        best[m] = k
    d = a[best] * xs + b[best] * ys + c[best]
    if fg_frac > 0:
        nfg = max(1, planes // 2)
        r = np.sqrt(fg_frac * W * H / (np.pi * nfg))
        for _ in range(nfg):
            cx, cy = rng.uniform(0, W), rng.uniform(0, H)
            disc = (xs - cx) ** 2 + (ys - cy) ** 2 < r * r
            if disc.any():
                near = min(dmax, float(d[disc].max()) + rng.uniform(2, 0.1 * dmax + 3))
                d[disc] = near
    return np.clip(np.rint(d), 0, dmax).astype(np.int32)


def warp_right(left, disp, maxval):
    """Forward warp left(x, y) -> right(x - d, y), larger d wins; holes take the background neighbour."""
    H, W = left.shape
    ys, xs = np.mgrid[0:H, 0:W]
    xr = xs - disp
    ok = xr >= 0
    key = np.full(H * W, -1, np.int64)
    tgt = (ys * W + xr)[ok]
    val = (disp[ok].astype(np.int64) << 20) | left[ok].astype(np.int64)
    np.maximum.at(key, tgt, val)
    key = key.reshape(H, W)
    hole = key < 0
    val = np.where(hole, 0, key & 0xFFFFF)
    dv = np.where(hole, np.iinfo(np.int64).max, key >> 20)
    idx = np.where(~hole, np.arange(W)[None, :], -1)
    lidx = np.maximum.accumulate(idx, axis=1)                         # nearest visible to the left
    ridx = np.where(~hole, np.arange(W)[None, :], W)
    ridx = np.minimum.accumulate(ridx[:, ::-1], axis=1)[:, ::-1]      # nearest visible to the right
    rowi = np.arange(H)[:, None]
    lv = np.where(lidx >= 0, val[rowi, np.clip(lidx, 0, W - 1)], 0)
    ld = np.where(lidx >= 0, dv[rowi, np.clip(lidx, 0, W - 1)], np.iinfo(np.int64).max)
    rv = np.where(ridx < W, val[rowi, np.clip(ridx, 0, W - 1)], 0)
    rd = np.where(ridx < W, dv[rowi, np.clip(ridx, 0, W - 1)], np.iinfo(np.int64).max)
    fill = np.where(ld <= rd, lv, rv)
    return np.where(hole, fill, val).astype(left.dtype)


def make_pair(W, H, dmax, *, planes=4, sigma=2.0, seed=0, bits=8, fg_frac=0.0):
    """Returns (left, right, disparity) as (uint8|uint16 [H, W], same, int32 [H, W])."""
    rng = np.random.default_rng(seed)
    maxval = (1 << bits) - 1
    dt = np.uint8 if bits == 8 else np.uint16
    noise = rng.integers(0, maxval + 1, (H, W)).astype(dt)
    left = box3(noise, maxval)
    disp = disparity_field(W, H, dmax, planes, rng, fg_frac)
    right = warp_right(left, disp, maxval).astype(np.float32)
    right = right + rng.standard_normal((H, W), dtype=np.float32) * np.float32(sigma)
    right = np.clip(np.rint(right), 0, maxval).astype(dt)
    return left, right, disp


def sha256(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def plane_list(near, far, n):
    """n metric plane depths evenly spaced in inverse depth, far .. near, the last one exactly `near` (the spacing of
    DenseStereo.Planes, DenseStereo.cs:41-88, without its reciprocal approximation)."""
    inv = np.linspace(1.0 / far, 1.0 / near, n)
    p = (1.0 / inv).astype(np.float32)
    p[-1] = np.float32(near)
    return p


def _margin(fbs, z_bg, z_fg):
    return int(max(abs(f) for f in fbs) / min(z_bg, z_fg)) + 8


def multiview_calibs(W0, H0, *, fbs=(4000.0, -6000.0, 8000.0, -2000.0), z_bg=2000.0, z_fg=1000.0, margin=None, f=1000.0):
    """CalibData records (K, T, R) of the rig `multiview` renders: focal length f, no rotation; the reference camera at
    the origin with its principal point at the image centre, camera k translated by -fbs[k] / f along x with its
    principal point `margin` further right. Mat4D.FlowMatrix of (reference, camera k) is then the warp `multiview`
    returns for camera k."""
    M = _margin(fbs, z_bg, z_fg) if margin is None else margin
    I = np.eye(3, dtype=np.float32)

    def K(cx):
        return np.array([[f, 0, cx], [0, f, H0 / 2], [0, 0, 1]], np.float32)

    cams = [(K(W0 / 2), np.zeros(3, np.float32), I)]
    for fb in fbs:
        cams.append((K(W0 / 2 + M), np.array([-fb / f, 0, 0], np.float32), I))
    return cams


def multiview(W0, H0, *, fbs=(4000.0, -6000.0, 8000.0, -2000.0), gray=(False, False, True, False), z_bg=2000.0,
              z_fg=1000.0, rects=2, seed=0):
    """A reference camera and len(fbs) rectified cameras looking at a textured fronto-parallel background at z_bg with
    `rects` nearer rectangles at z_fg. Camera k sees reference pixel x at x - fbs[k]/z + M (M = a margin that keeps
    every projection of the planes inside its wider image); its warp is the flow matrix of that map (column-major:
    m0 = m5 = m10 = m15 = 1, m8 = M, m12 = -fb). Images are RGBA8 noise (numpy PCG64, `seed`), not yet box-filtered;
    a gray camera's image repeats channel 0. Returns (images, gray flags incl. the reference, warps [k][16], depth)."""
    rng = np.random.default_rng(seed)
    M = _margin(fbs, z_bg, z_fg)
    Wt = W0 + 2 * M
    bg = rng.integers(0, 256, (H0, Wt, 4), dtype=np.uint8)
    fg = rng.integers(0, 256, (H0, W0, 4), dtype=np.uint8)
    depth = np.full((H0, W0), np.float32(z_bg), np.float32)
    ref = bg[:, M:M + W0].copy()
    boxes = []
    for _ in range(rects):
        w, h = int(rng.integers(W0 // 6, W0 // 3)), int(rng.integers(H0 // 6, H0 // 3))
        x0, y0 = int(rng.integers(M, W0 - w - M)), int(rng.integers(0, H0 - h))
        boxes.append((x0, y0, w, h))
        ref[y0:y0 + h, x0:x0 + w] = fg[y0:y0 + h, x0:x0 + w]
        depth[y0:y0 + h, x0:x0 + w] = z_fg
    images, warps = [ref], []
    for fb, g in zip(fbs, gray):
        sb, sf = int(round(fb / z_bg)), int(round(fb / z_fg))
        img = np.empty((H0, Wt, 4), np.uint8)
        # background: camera pixel x' shows reference-plane pixel x = x' + sb - M of the wide texture
        xs = np.clip(np.arange(Wt) + sb, 0, Wt - 1)
        img[:] = bg[:, xs]
        for x0, y0, w, h in boxes:
            img[y0:y0 + h, x0 - sf + M:x0 - sf + M + w] = fg[y0:y0 + h, x0:x0 + w]
        if g:
            img[:, :, 1] = img[:, :, 2] = img[:, :, 3] = img[:, :, 0]
        images.append(img)
        m = np.zeros(16, np.float32)
        m[0] = m[5] = m[10] = m[15] = 1.0
        m[8] = M
        m[12] = -fb
        warps.append(m)
    return images, [False] + list(gray), warps, depth
