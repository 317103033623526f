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
