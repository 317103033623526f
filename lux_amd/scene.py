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
