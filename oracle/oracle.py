"""ctypes wrapper of the CPU oracle (liblux_oracle.so).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()`` as the checker and by
``bench.py``'s CPU-baseline leg. The product path (``lux_amd``) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblux_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        L.lo_census.argtypes = [vp, i, i, i, i, i, vp]
        L.lo_cost_census.argtypes = [vp, vp, i, i, i, i, vp, i]
        L.lo_cost_ad.argtypes = [vp, vp, i, i, i, i, i, i, i, i, vp, i]
        L.lo_box3.argtypes = [vp, i, i, i, i, vp]
        L.lo_layer_compute.argtypes = [vp, vp, vp, vp, vp]
        L.lo_wta_sub_lr.argtypes = [vp, i, i, i, i, i, vp, vp, vp]
        L.lo_p2_adaptive.argtypes = [f, vp, vp, vp]
        L.lo_set_force_scalar.argtypes = [i]
        L.lo_rectified.argtypes = [vp, vp, i, i, i, i, i, i, i, i, i, i, i, i, i, i, i, i, vp, vp, vp, vp, vp]
        _LIB = L
    return _LIB


class LayerParams(ctypes.Structure):
    _fields_ = [
        ("W", ctypes.c_int), ("H", ctypes.c_int), ("T", ctypes.c_int), ("align", ctypes.c_int), ("xsw", ctypes.c_int),
        ("p1", ctypes.c_int), ("p2_const", ctypes.c_int), ("p2_scale", ctypes.c_float), ("gw", ctypes.c_float * 4),
        ("guidance", ctypes.c_void_p), ("start", ctypes.c_void_p), ("count", ctypes.c_void_p), ("skip", ctypes.c_void_p),
    ]


def _p(a):
    return None if a is None else a.ctypes.data


def census(img, radius):
    img = np.ascontiguousarray(img)
    H, W = img.shape
    out = np.empty((H, W), np.uint64)
    rc = lib().lo_census(_p(img), int(img.dtype == np.uint16), W, H, W, radius, _p(out))
    assert rc == 0, rc
    return out


def cost_census(cl, cr, D, cap, threads=1):
    H, W = cl.shape
    raw = np.empty((H, W, cap), np.uint8)
    rc = lib().lo_cost_census(_p(np.ascontiguousarray(cl)), _p(np.ascontiguousarray(cr)), W, H, D, cap, _p(raw), threads)
    assert rc == 0, rc
    return raw


def cost_ad(left, right, bits, D, cap, adcap=28, threads=1):
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    H, W = left.shape
    raw = np.empty((H, W, cap), np.uint8)
    rc = lib().lo_cost_ad(_p(left), _p(right), int(left.dtype == np.uint16), bits, W, H, W, D, cap, adcap, _p(raw), threads)
    assert rc == 0, rc
    return raw


def box3(img, maxval):
    img = np.ascontiguousarray(img)
    H, W = img.shape
    out = np.empty_like(img)
    rc = lib().lo_box3(_p(img), int(img.dtype == np.uint16), W, H, maxval, _p(out))
    assert rc == 0, rc
    return out


def layer(raw_packed, W, H, start, count, T=64, align=8, p1=10, p2=120, xsw=False, skip=None, guidance=None,
          gw=(0.0, 0.0, 0.0, 0.0), p2_scale=0.0, adaptive=False, keep_pass1=False):
    """DenseLayer.Compute restated. raw_packed: uint8 flat, cap entries per pixel (raster order).
    Returns (agg_packed uint16 flat, agg_pass1 or None, wta uint16 [H, W])."""
    start = np.ascontiguousarray(start, np.uint16).reshape(-1)
    count = np.ascontiguousarray(count, np.uint16).reshape(-1)
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    n = int(caps.sum())
    raw_packed = np.ascontiguousarray(raw_packed, np.uint8).reshape(-1)
    assert raw_packed.size == n, (raw_packed.size, n)
    agg = np.empty(n, np.uint16)
    agg1 = np.empty(n, np.uint16) if keep_pass1 else None
    wta = np.empty(W * H, np.uint16)
    lp = LayerParams()
    lp.W, lp.H, lp.T, lp.align, lp.xsw, lp.p1 = W, H, T, align, int(bool(xsw)), p1
    lp.p2_const = -1 if adaptive else p2
    lp.p2_scale = p2_scale
    for k in range(4):
        lp.gw[k] = gw[k]
    g = None if guidance is None else np.ascontiguousarray(guidance, np.uint8)
    sk = None if skip is None else np.ascontiguousarray(skip, np.uint8).reshape(-1)
    lp.guidance, lp.start, lp.count, lp.skip = _p(g), _p(start), _p(count), _p(sk)
    rc = lib().lo_layer_compute(ctypes.byref(lp), _p(raw_packed), _p(agg), _p(agg1), _p(wta))
    if rc != 0:
        raise RuntimeError(f"lo_layer_compute failed: {rc}")
    return agg, agg1, wta.reshape(H, W)


def wta_sub_lr(agg, D, tol=1):
    agg = np.ascontiguousarray(agg, np.uint16)
    H, W, cap = agg.shape
    disp = np.empty((H, W), np.uint16)
    sub = np.empty((H, W), np.float32)
    valid = np.empty((H, W), np.uint8)
    rc = lib().lo_wta_sub_lr(_p(agg), W, H, D, cap, tol, _p(disp), _p(sub), _p(valid))
    assert rc == 0, rc
    return disp, sub, valid


def rectified(left, right, D, *, cost="census", radius=3, bits=10, adcap=28, align=8, p1=8, p2=96, T=64, xsw=False,
              tol=1, threads=1, want_agg=False):
    """Whole CPU pipeline. Returns dict(disp, sub, valid, agg?, times_ms)."""
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    H, W = left.shape
    cap = ((D + align - 1) // align) * align
    disp = np.empty((H, W), np.uint16)
    sub = np.empty((H, W), np.float32)
    valid = np.empty((H, W), np.uint8)
    agg = np.empty((H, W, cap), np.uint16) if want_agg else None
    times = np.zeros(3, np.float64)
    rc = lib().lo_rectified(_p(left), _p(right), int(left.dtype == np.uint16), bits, W, H, W,
                            0 if cost == "census" else 1, radius, adcap, D, align, p1, p2, T, int(bool(xsw)), tol,
                            threads, _p(disp), _p(sub), _p(valid), _p(agg), _p(times))
    if rc != 0:
        raise RuntimeError(f"lo_rectified failed: {rc}")
    out = dict(disp=disp, sub=sub, valid=valid, times_ms=times)
    if want_agg:
        out["agg"] = agg
    return out


def downsample2(img):
    """2x2 rounded mean, the last row / column repeated on odd sizes (one coarse-to-fine level)."""
    H, W = img.shape
    H2, W2 = (H + 1) // 2, (W + 1) // 2
    ys0 = np.minimum(2 * np.arange(H2), H - 1)
    ys1 = np.minimum(2 * np.arange(H2) + 1, H - 1)
    xs0 = np.minimum(2 * np.arange(W2), W - 1)
    xs1 = np.minimum(2 * np.arange(W2) + 1, W - 1)
    a = img.astype(np.uint32)
    s = a[ys0][:, xs0] + a[ys0][:, xs1] + a[ys1][:, xs0] + a[ys1][:, xs1]
    return ((s + 2) >> 2).astype(img.dtype)


def cost_census_ranges(cl, cr, start, n):
    """raw[y, x, j] = popcount(cl(x) ^ cr(max(x - (start + j), 0))) for j < n (the census cost of a range)."""
    H, W = cl.shape
    xr = np.maximum(np.arange(W)[None, :, None] - (start.astype(np.int64)[:, :, None] + np.arange(n)[None, None, :]), 0)
    crs = cr[np.arange(H)[:, None, None], xr]
    x = np.bitwise_xor(cl[:, :, None], crs).view(np.uint8).reshape(H, W, n, 8)
    return np.unpackbits(x, axis=3).sum(axis=3).astype(np.uint8)


def pyramid(left, right, D, radii, *, radius=3, bits=10, align=8, p1=8, p2=96, T=64):
    """Coarse-to-fine rectified request from the oracle's parts: the coarsest level as `rectified`, every finer
    level as `layer` over per-pixel ranges start = clip(2 d - r, 0, D_k - 2r), count = 2r, with the census
    cost of those planes. Returns (disp, sub) of the finest level."""
    levels = len(radii) + 1
    imgs = [(np.ascontiguousarray(left), np.ascontiguousarray(right))]
    for _ in range(levels - 1):
        imgs.insert(0, (downsample2(imgs[0][0]), downsample2(imgs[0][1])))
    l0, r0 = imgs[0]
    out = rectified(l0, r0, D >> (levels - 1), cost="census", radius=radius, bits=bits, align=align, p1=p1, p2=p2, T=T)
    disp, sub = out["disp"], out["sub"]
    for k in range(1, levels):
        lk, rk = imgs[k]
        H, W = lk.shape
        Dk, r = D >> (levels - 1 - k), radii[k - 1]
        yc = np.minimum(np.arange(H) // 2, disp.shape[0] - 1)
        xc = np.minimum(np.arange(W) // 2, disp.shape[1] - 1)
        dprev = disp.astype(np.int64)[yc][:, xc]
        start = np.clip(2 * dprev - r, 0, Dk - 2 * r).astype(np.int64)
        n = 2 * r
        raw = cost_census_ranges(census(lk, radius), census(rk, radius), start, n)
        count = np.full((H, W), n, np.uint16)
        agg, _, wta = layer(raw.reshape(-1), W, H, start.astype(np.uint16), count, T=T, align=align, p1=p1, p2=p2)
        agg = agg.reshape(H, W, n).astype(np.int64)
        arg = wta.astype(np.int64) - start
        disp = wta
        sub = (start + arg).astype(np.float32)
        a = np.take_along_axis(agg, np.clip(arg - 1, 0, n - 1)[..., None], 2)[..., 0]
        c = np.take_along_axis(agg, arg[..., None], 2)[..., 0]
        b = np.take_along_axis(agg, np.clip(arg + 1, 0, n - 1)[..., None], 2)[..., 0]
        den = a - 2 * c + b
        ok = (arg > 0) & (arg < n - 1) & (den > 0)
        off = (np.float32(0.5) * (a - b).astype(np.float32)) / np.where(ok, den, 1).astype(np.float32)
        sub = np.where(ok, sub + off.astype(np.float32), sub).astype(np.float32)
    return disp, sub
