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
        L.lo_layer_compute_visits.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.c_int32]
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
          gw=(0.0, 0.0, 0.0, 0.0), p2_scale=0.0, adaptive=False, keep_pass1=False, visits=None, n_plus=0):
    """DenseLayer.Compute restated. raw_packed: uint8 flat, cap entries per pixel (raster order).
    visits (optional, int32 [n, 3] of x, y, cur half; the first n_plus in pass +1): run Pixel in that order
    instead of the reference's tile loop. Returns (agg_packed uint16 flat, agg_pass1 or None, wta uint16 [H, W])."""
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
    if visits is None:
        rc = lib().lo_layer_compute(ctypes.byref(lp), _p(raw_packed), _p(agg), _p(agg1), _p(wta))
    else:
        v = np.ascontiguousarray(visits, np.int32).reshape(-1, 3)
        rc = lib().lo_layer_compute_visits(ctypes.byref(lp), _p(raw_packed), _p(agg), _p(agg1), _p(wta), _p(v),
                                           int(n_plus), int(v.shape[0] - n_plus))
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


def _upsample_lib():
    L = lib()
    if not getattr(L, "_ups", False):
        vp, i32, u32, f32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_float
        L.lo_rcp_bits.argtypes, L.lo_rcp_bits.restype = [u32], u32
        L.lo_rcpss_hw.argtypes, L.lo_rcpss_hw.restype = [u32], u32
        L.lo_image_shift.argtypes = [vp, i32, i32, f32, f32, vp]
        L.lo_nearest_resize.argtypes = [vp, i32, i32, i32, i32, vp]
        L.lo_bilateral_upsample.argtypes = [vp, i32, i32, vp, i32, i32, vp, f32, vp]
        L.lo_upsample_layer.argtypes = [vp, i32, i32, vp, i32, i32, i32, i32, f32, vp]
        L.lo_rsqrt_bits.argtypes, L.lo_rsqrt_bits.restype = [u32], u32
        L.lo_rsqrtss_hw.argtypes, L.lo_rsqrtss_hw.restype = [u32], u32
        L.lo_approx_mismatches.argtypes, L.lo_approx_mismatches.restype = [i32, ctypes.c_uint64, ctypes.c_uint64], \
            ctypes.c_uint64
        L.lo_homography_sqrt.argtypes, L.lo_homography_sqrt.restype = [f32], f32
        L.lo_planes.argtypes, L.lo_planes.restype = [f32, f32, vp, i32, f32, i32, vp, i32, vp], i32
        L.lo_flow_matrix.argtypes = [vp, vp, vp]
        L._ups = True
    return L


def rcp_bits(b):
    return int(_upsample_lib().lo_rcp_bits(int(b) & 0xFFFFFFFF))


def rcpss_hw(b):
    return int(_upsample_lib().lo_rcpss_hw(int(b) & 0xFFFFFFFF))


def rsqrt_bits(b):
    return int(_upsample_lib().lo_rsqrt_bits(int(b) & 0xFFFFFFFF))


def rsqrtss_hw(b):
    return int(_upsample_lib().lo_rsqrtss_hw(int(b) & 0xFFFFFFFF))


def approx_mismatches(which, lo, hi):
    """How many bit patterns in [lo, hi) the rcpss (which="rcp") / rsqrtss ("rsqrt") model maps differently from the
    local CPU's instruction."""
    return int(_upsample_lib().lo_approx_mismatches(1 if which == "rsqrt" else 0, int(lo), int(hi)))


def homography_sqrt(x):
    return np.float32(_upsample_lib().lo_homography_sqrt(float(np.float32(x))))


def _calib_records(calibs):
    """calibs: sequence of (K 3x3, T 3, R 3x3) -> float32 [n, 21] records K, T, R (row-major)."""
    rec = np.empty((len(calibs), 21), np.float32)
    for k, (K, T, R) in enumerate(calibs):
        rec[k, :9] = np.asarray(K, np.float32).reshape(9)
        rec[k, 9:12] = np.asarray(T, np.float32).reshape(3)
        rec[k, 12:] = np.asarray(R, np.float32).reshape(9)
    return rec


PLANES_ERRORS = {-1: "near >= far", -2: "non-positive depth", -3: "baseline too small", -4: "bad arguments"}


def planes(near_z, far_z, calibs, density, align=1):
    """DenseStereo.Planes (DenseStereo.cs:41-88): the float32 plane depths, far .. near. Raises ValueError with the
    reason where the reference throws."""
    rec = _calib_records(calibs)
    n = ctypes.c_int(0)
    L = _upsample_lib()
    rc = L.lo_planes(float(near_z), float(far_z), _p(rec), len(rec), float(density), int(align), None, 0,
                     ctypes.byref(n))
    if rc:
        raise ValueError(PLANES_ERRORS[rc])
    out = np.empty(n.value, np.float32)
    rc = L.lo_planes(float(near_z), float(far_z), _p(rec), len(rec), float(density), int(align), _p(out), n.value,
                     ctypes.byref(n))
    if rc:
        raise ValueError(PLANES_ERRORS[rc])
    return out


def flow_matrix(a, b):
    """Mat4D.FlowMatrix (ViewTransform.cs:120-133) of calibrations a, b ((K, T, R) each): column-major float32[16]."""
    ra, rb = _calib_records([a]), _calib_records([b])
    out = np.empty(16, np.float32)
    _upsample_lib().lo_flow_matrix(_p(ra), _p(rb), _p(out))
    return out


def image_shift(img, sx, sy):
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape[:2]
    out = np.empty_like(img)
    assert _upsample_lib().lo_image_shift(_p(img), W, H, sx, sy, _p(out)) == 0
    return out


def nearest_resize(img, W, H):
    img = np.ascontiguousarray(img, np.uint8)
    out = np.empty((H, W, 4), np.uint8)
    assert _upsample_lib().lo_nearest_resize(_p(img), img.shape[1], img.shape[0], W, H, _p(out)) == 0
    return out


def bilateral_upsample(lo_img, guide, g2, sigma):
    lo_img = np.ascontiguousarray(lo_img, np.float32)
    guide = np.ascontiguousarray(guide, np.uint8)
    g2 = np.ascontiguousarray(g2, np.uint8)
    H, W = guide.shape[:2]
    out = np.empty((H, W), np.float32)
    assert _upsample_lib().lo_bilateral_upsample(_p(lo_img), lo_img.shape[1], lo_img.shape[0], _p(guide), W, H, _p(g2),
                                                 sigma, _p(out)) == 0
    return out


def upsample_layer(prev_depth, guide, set_w, set_h, sigma=12.0):
    """DenseUpsampleLayer.Run (DenseUpsampleLayer.cs:38-52): the finest layer's depth to the guide's resolution."""
    prev_depth = np.ascontiguousarray(prev_depth, np.float32)
    guide = np.ascontiguousarray(guide, np.uint8)
    H, W = guide.shape[:2]
    out = np.empty((H, W), np.float32)
    rc = _upsample_lib().lo_upsample_layer(_p(prev_depth), prev_depth.shape[1], prev_depth.shape[0], _p(guide), W, H,
                                           set_w, set_h, sigma, _p(out))
    if rc != 0:
        raise RuntimeError(f"lo_upsample_layer failed: {rc}")
    return out


# ---------------------------------------------------------------- plane-sweep matching cost (DenseStereo.cs)
class View(ctypes.Structure):
    _fields_ = [("rgba", ctypes.c_void_p), ("w", ctypes.c_int), ("h", ctypes.c_int), ("warp", ctypes.c_float * 16),
                ("wq", ctypes.c_uint16 * 4)]


class Sweep(ctypes.Structure):
    _fields_ = [("ref", ctypes.c_void_p), ("rw", ctypes.c_int), ("rh", ctypes.c_int), ("views", ctypes.c_void_p),
                ("nv", ctypes.c_int), ("planes", ctypes.c_void_p), ("n", ctypes.c_int), ("caps", ctypes.c_uint8 * 4)]


def _sweep_lib():
    L = _upsample_lib()
    if not getattr(L, "_sweep", False):
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        L.lo_matching_cost.argtypes = [vp, i32, i32, i32, i32, vp]
        L.lo_normalise.argtypes = [vp, i32, i32, i32, vp]
        L.lo_normalise.restype = None
        L.lo_layer_sweep_cost.argtypes = [vp, i32, i32, i32, i32, vp, vp, vp, vp, i32]
        L.lo_guidance.argtypes = [vp, i32, i32, i32, i32, i32, vp]
        L.lo_range_from_previous.argtypes = [vp, i32, vp, vp, i32, i32, i32, i32, i32, vp, vp, vp, vp]
        L.lo_layer_depth.argtypes = [vp, i32, i32, i32, i32, vp, vp, vp, vp]
        L.lo_skip_mask.argtypes = [i32, i32, i32, vp]
        L.lo_mt19937.argtypes, L.lo_mt19937.restype = [ctypes.c_uint32, i32], ctypes.c_uint32
        L._sweep = True
    return L


class SweepContext:
    """The cost context of a layer (DenseStereo.CostContext): the filtered reference, the other cameras' filtered
    images, warps (16 floats, column-major) and weights, the planes and the caps. Keeps its arrays alive."""

    def __init__(self, ref, others, warps, wqs, planes, caps=(2, 6, 6, 0)):
        self.ref = np.ascontiguousarray(ref, np.uint8)
        self.others = [np.ascontiguousarray(o, np.uint8) for o in others]
        self.planes = np.ascontiguousarray(planes, np.float32)
        self.views = (View * len(others))()
        for k, (img, w, q) in enumerate(zip(self.others, warps, wqs)):
            self.views[k].rgba = img.ctypes.data
            self.views[k].h, self.views[k].w = img.shape[:2]
            self.views[k].warp = (ctypes.c_float * 16)(*np.asarray(w, np.float32).reshape(16))
            self.views[k].wq = (ctypes.c_uint16 * 4)(*[int(v) for v in q])
        self.s = Sweep(self.ref.ctypes.data, self.ref.shape[1], self.ref.shape[0], ctypes.addressof(self.views),
                       len(others), self.planes.ctypes.data, len(self.planes), (ctypes.c_uint8 * 4)(*caps))


def matching_cost(sc, x, y, start, count):
    out = np.zeros(max(count, 1), np.uint16)
    rc = _sweep_lib().lo_matching_cost(ctypes.byref(sc.s), x, y, start, count, _p(out))
    assert rc == 0, rc
    return out[:count]


def normalise(cost, count, cap, nv):
    cost = np.ascontiguousarray(cost, np.uint16)
    raw = np.zeros(max(cap, 1), np.uint8)
    _sweep_lib().lo_normalise(_p(cost), count, cap, nv, _p(raw))
    return raw[:cap]


def layer_sweep_cost(sc, W, H, scale, align, start, count, skip=None, threads=8):
    """The packed raw volume of a layer: roundup(count, align) bytes per pixel in raster order."""
    start = np.ascontiguousarray(start, np.uint16).reshape(-1)
    count = np.ascontiguousarray(count, np.uint16).reshape(-1)
    skip = None if skip is None else np.ascontiguousarray(skip, np.uint8).reshape(-1)
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    raw = np.zeros(max(int(caps.sum()), 1), np.uint8)
    rc = _sweep_lib().lo_layer_sweep_cost(ctypes.byref(sc.s), W, H, scale, align, _p(start), _p(count), _p(skip), _p(raw),
                                          threads)
    assert rc == 0, rc
    return raw[:int(caps.sum())]


def guidance(ref, w, h, scale):
    ref = np.ascontiguousarray(ref, np.uint8)
    out = np.empty((h, w, 4), np.uint8)
    assert _sweep_lib().lo_guidance(_p(ref), ref.shape[1], ref.shape[0], w, h, scale, _p(out)) == 0
    return out


def range_from_previous(planes, prev_depth, prev_skip, w, h, margin=1):
    planes = np.ascontiguousarray(planes, np.float32)
    prev_depth = np.ascontiguousarray(prev_depth, np.float32)
    ph, pw = prev_depth.shape
    prev_skip = None if prev_skip is None else np.ascontiguousarray(prev_skip, np.uint8)
    start = np.empty((h, w), np.uint16)
    count = np.empty((h, w), np.uint16)
    lo, hi = ctypes.c_int(), ctypes.c_int()
    rc = _sweep_lib().lo_range_from_previous(_p(planes), len(planes), _p(prev_depth), _p(prev_skip), pw, ph, w, h, margin,
                                             _p(start), _p(count), ctypes.byref(lo), ctypes.byref(hi))
    assert rc == 0, rc
    return start, count, lo.value, hi.value


def layer_depth(planes, W, H, align, start, count, agg):
    planes = np.ascontiguousarray(planes, np.float32)
    start = np.ascontiguousarray(start, np.uint16).reshape(-1)
    count = np.ascontiguousarray(count, np.uint16).reshape(-1)
    agg = np.ascontiguousarray(agg, np.uint16).reshape(-1)
    out = np.empty((H, W), np.float32)
    assert _sweep_lib().lo_layer_depth(_p(planes), len(planes), W, H, align, _p(start), _p(count), _p(agg), _p(out)) == 0
    return out


def dense_layer(sc, W, H, scale, tile, align, start=None, count=None, skip=None, p1=1, p2_scale=500.0, gw=(0, 0, 0, 0),
                keep_pass1=False, threads=8):
    """DenseLayer.Compute from the images: plane-sweep cost, adaptive eight-path aggregation over the subsampled
    reference as guidance, WTA, depth. Returns dict(depth, wta, agg, agg_pass1, raw)."""
    n = len(sc.planes)
    st = np.zeros((H, W), np.uint16) if start is None else np.asarray(start, np.uint16).reshape(H, W)
    cn = np.full((H, W), n, np.uint16) if count is None else np.asarray(count, np.uint16).reshape(H, W)
    raw = layer_sweep_cost(sc, W, H, scale, align, st, cn, skip, threads=threads)
    g = guidance(sc.ref, W, H, scale)
    agg, agg1, wta = layer(raw, W, H, st, cn, T=tile, align=align, p1=p1, skip=skip, guidance=g, gw=gw, p2_scale=p2_scale,
                           adaptive=True, keep_pass1=keep_pass1)
    depth = layer_depth(sc.planes, W, H, align, st, cn, agg)
    return dict(depth=depth, wta=wta, agg=agg, agg_pass1=agg1, raw=raw)


def skip_mask(mode, w, h):
    out = np.empty((h, w), np.uint8)
    rc = _sweep_lib().lo_skip_mask(mode, w, h, _p(out))
    assert rc == 0, rc
    return out


def mt19937(seed, n):
    return int(_sweep_lib().lo_mt19937(seed, n))


def box3_rgba(img):
    """BoxFilter3 of an RGBA8 image: the single-channel filter on each channel (the sums never mix channels)."""
    return np.stack([box3(np.ascontiguousarray(img[:, :, c]), 255) for c in range(4)], axis=2)


def memory_estimate(W, H, start, count, skip, align, maxhi, image_bytes):
    """DenseLayer.MemoryEstimate (DenseStereoSgm.cs:142-156) from the layer's ranges (before rounding MaxHi)."""
    cap = ((np.asarray(count, np.int64).reshape(-1) + align - 1) // align) * align
    sk = np.asarray(skip).reshape(-1) != 0
    cost = int((8 + cap * np.where(sk, 2, 3)).sum())
    mh = -(-maxhi // align) * align
    w, h = W, H
    return cost + w * h * 4 + 20 * (2 * w + 4) + 2 * ((4 * mh + 5) * (2 * w + 4) + 8 * w + 16) + image_bytes + w * h * 8


L16 = dict(scales=(32, 16, 8, 4, 2, 1), tiles=(4, 8, 16, 32, 64, 64), skip_modes=(0, 0, 0, 0, 2, 2))


def dense_pyramid(images, gray, warps, planes8, planes2=None, scales=L16["scales"], tiles=L16["tiles"],
                  skip_modes=L16["skip_modes"], p1=1, p2_scale=500.0, gw=(0, 0, 0, 0), margin=1, budget=1 << 30,
                  threads=8):
    """DenseStereoPyramid.Run (DenseStereoSgm.cs:331-355) composed from the oracle's parts. Returns a list of per-level
    dicts (depth, align, start, count, skip, minlo, maxhi)."""
    filtered = [box3_rgba(np.ascontiguousarray(im, np.uint8)) for im in images]
    H0, W0 = filtered[0].shape[:2]
    image_bytes = sum(f.shape[0] * f.shape[1] * 4 for f in filtered)
    wqs = [(12240, 0, 0, 0) if (gray is not None and gray[k]) else (8160, 680, 680, 0) for k in range(1, len(images))]
    out = []
    prev = None
    for i, (sc_, tile, sm) in enumerate(zip(scales, tiles, skip_modes)):
        W, H = -(-W0 // sc_), -(-H0 // sc_)
        skip = skip_mask(sm, W, H)
        align, planes = 8, np.asarray(planes8, np.float32)
        start = count = None
        minlo, maxhi = 0, len(planes)
        if prev is not None:
            start, count, minlo, maxhi = range_from_previous(planes8, prev["depth"], prev["skip"], W, H, margin)
            if memory_estimate(W, H, start, count, skip, 8, maxhi, image_bytes) >= budget:
                if planes2 is None:
                    raise RuntimeError("align-2 planes needed")
                start, count, minlo, maxhi = range_from_previous(planes2, prev["depth"], prev["skip"], W, H, margin)
                if memory_estimate(W, H, start, count, skip, 2, maxhi, image_bytes) >= budget:
                    raise RuntimeError("memory budget exceeded")
                align, planes = 2, np.asarray(planes2, np.float32)
        sc = SweepContext(filtered[0], filtered[1:], warps, wqs, planes)
        r = dense_layer(sc, W, H, sc_, tile, align, start, count, skip, p1=p1, p2_scale=p2_scale, gw=gw, threads=threads)
        lvl = dict(depth=r["depth"], align=align, start=start, count=count, skip=skip, minlo=minlo, maxhi=maxhi, W=W, H=H)
        out.append(lvl)
        prev = lvl
    return out
