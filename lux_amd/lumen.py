"""Lumen's dense-stereo dump format as the fixture / wire format of a layer, and the ``depthOverride`` seam.

Formats (Lux.Engine/Pipeline/Registration/StereoAsyncApi.cs):
  image files (``DumpRgba8`` :212-219, ``DumpF32`` :249-, the ``_skip.u8`` dump :196): int32 {w, h, stride (pixels),
      bytes per pixel} then ``h`` rows of ``w * bpp`` bytes;
  ``<prefix>_offsets.u32`` (``DumpCostVolume`` :224-248): the same 16-byte header (bpp 4) then one u32 byte offset per
      pixel into the cost file;
  ``<prefix>_costvol.bin`` / ``_costvol_pass1.bin``: per pixel an 8-byte header {start u16, count u16, 1 u16, cap u16},
      ``cap`` u16 aggregates, then (computed pixels only) ``cap`` u8 raw costs.
The depth seam: ``StereoAsyncApi.Run(..., depthOverride)`` (:48-63) takes the reference camera's depth as a float
array and uses it in place of the dense stage (:207), so a depth map from the device path swaps in with no API change.
"""
import os

import numpy as np

from . import sgm


def write_image(path, arr):
    """Write a [h, w] or [h, w, c] array in the dump image format (bpp = itemsize * channels)."""
    a = np.ascontiguousarray(arr)
    h, w = a.shape[:2]
    bpp = a.itemsize * (a.shape[2] if a.ndim == 3 else 1)
    with open(path, "wb") as f:
        f.write(np.array([w, h, w, bpp], "<i4").tobytes())
        f.write(a.tobytes())


def read_image(path, dtype):
    """Read a dump image; dtype is the element type (uint8 for RGBA8 / skip, float32 for depth)."""
    b = np.fromfile(path, np.uint8)
    w, h, stride, bpp = (int(v) for v in b[:16].view("<i4"))
    dt = np.dtype(dtype)
    ch = bpp // dt.itemsize
    rows = b[16:16 + h * stride * bpp].reshape(h, stride * bpp)[:, : w * bpp]
    a = np.ascontiguousarray(rows).view(dt).reshape(h, w, ch) if ch > 1 else np.ascontiguousarray(rows).view(dt).reshape(h, w)
    return a


def caps_of(count, align):
    return ((np.asarray(count, np.int64) + align - 1) // align) * align


def write_costvol(prefix, W, H, start, count, align, agg, raw=None, skip=None, agg_pass1=None):
    """Write a layer as Lumen does: ``agg`` / ``raw`` jagged (cap_i entries per pixel, raster order) or [H, W, cap]."""
    start = np.asarray(start, np.uint16).reshape(-1)
    count = np.asarray(count, np.uint16).reshape(-1)
    cap = caps_of(count, align)
    skip = np.zeros(W * H, np.uint8) if skip is None else np.asarray(skip, np.uint8).reshape(-1)
    has_raw = (skip == 0) & (raw is not None)
    size = 8 + cap * np.where(has_raw, 3, 2)
    off = np.concatenate(([0], np.cumsum(size)[:-1])).astype(np.uint64)
    assert int(off[-1] + size[-1]) < 2 ** 32, "the offsets are u32"
    hdr = np.array([W, H, W, 4], "<i4").tobytes()
    with open(prefix + "_offsets.u32", "wb") as f:
        f.write(hdr)
        f.write(off.astype("<u4").tobytes())
    coff = np.concatenate(([0], np.cumsum(cap)[:-1])).astype(np.int64)

    def one(path, a):
        a = np.asarray(a, np.uint16).reshape(-1)
        r = None if raw is None else np.asarray(raw, np.uint8).reshape(-1)
        buf = bytearray(int(off[-1] + size[-1]))
        for i in range(W * H):
            o, c, k = int(off[i]), int(cap[i]), int(coff[i])
            buf[o:o + 8] = np.array([start[i], count[i], 1, c], "<u2").tobytes()
            buf[o + 8:o + 8 + 2 * c] = a[k:k + c].astype("<u2").tobytes()
            if has_raw[i]:
                buf[o + 8 + 2 * c:o + 8 + 3 * c] = r[k:k + c].tobytes()
        with open(path, "wb") as f:
            f.write(bytes(buf))

    one(prefix + "_costvol.bin", agg)
    if agg_pass1 is not None:
        one(prefix + "_costvol_pass1.bin", agg_pass1)


def read_costvol(prefix, pass1=False):
    """Read a layer dump: dict(W, H, start, count, cap, skip, agg, raw) with jagged agg / raw (raw 0 / 255 for
    pixels written without costs, which are the skipped ones)."""
    ob = np.fromfile(prefix + "_offsets.u32", np.uint8)
    W, H, _, _ = (int(v) for v in ob[:16].view("<i4"))
    off = ob[16:16 + 4 * W * H].view("<u4").astype(np.int64)
    buf = np.fromfile(prefix + ("_costvol_pass1.bin" if pass1 else "_costvol.bin"), np.uint8)
    n = W * H
    start = np.empty(n, np.uint16)
    count = np.empty(n, np.uint16)
    cap = np.empty(n, np.int64)
    ends = np.append(off[1:], buf.size)
    skip = np.zeros(n, np.uint8)
    aggs, raws = [], []
    for i in range(n):
        o = int(off[i])
        s_, c_, _, k = (int(v) for v in buf[o:o + 8].view("<u2"))
        start[i], count[i], cap[i] = s_, c_, k
        aggs.append(buf[o + 8:o + 8 + 2 * k].view("<u2").astype(np.uint16))
        if int(ends[i]) - o >= 8 + 3 * k:
            raws.append(buf[o + 8 + 2 * k:o + 8 + 3 * k].copy())
        else:  # a skipped pixel: 0 over its planes, 255 padding (DenseStereoSgm.cs:242)
            skip[i] = 0xFF
            r = np.full(k, 255, np.uint8)
            r[:c_] = 0
            raws.append(r)
    return dict(W=W, H=H, start=start, count=count, cap=cap.astype(np.uint16), skip=skip,
                agg=np.concatenate(aggs) if aggs else np.zeros(0, np.uint16),
                raw=np.concatenate(raws) if raws else np.zeros(0, np.uint8))


def layer_from_dump(ctx, prefix, planes, *, align=8, p1=1, p2=0, tile=4):
    """A DenseLayer whose ranges, skip mask and raw costs come from a Lumen cost-volume dump."""
    d = read_costvol(prefix)
    L = sgm.DenseLayer(ctx, d["W"], d["H"], planes, align=align, p1=p1, p2=p2, tile=tile)
    L.set_ranges(d["start"], d["count"])
    L.Skip = d["skip"]
    L.Raw = d["raw"]
    return L


def dump_layer(prefix, L):
    """Write a computed DenseLayer in Lumen's format (with the pass-1 aggregates when kept)."""
    start = getattr(L, "Start")
    count = getattr(L, "Count")
    write_costvol(prefix, L.W, L.H, start, count, L.Align, np.asarray(L.Agg).reshape(-1),
                  raw=np.asarray(L.Raw).reshape(-1), skip=getattr(L, "Skip", None),
                  agg_pass1=None if L.AggPass1 is None else np.asarray(L.AggPass1).reshape(-1))


def depth_override(sub, valid, focal_baseline, invalid=0.0):
    """The float[] depthOverride of StereoAsyncApi.Run from a rectified result: depth = f*B / disparity (float32)
    where the left/right check passed and the disparity is positive, `invalid` elsewhere; row-major W*H."""
    s = np.asarray(sub, np.float32)
    ok = (np.asarray(valid) != 0) & (s > 0)
    depth = np.full(s.shape, np.float32(invalid), np.float32)
    depth[ok] = np.float32(focal_baseline) / s[ok]
    return depth.reshape(-1)


def write_depth(path, depth, W, H):
    """``DumpF32``: the depth map in the dump image format."""
    write_image(path, np.asarray(depth, np.float32).reshape(H, W))


def exists(prefix):
    return os.path.exists(prefix + "_offsets.u32") and os.path.exists(prefix + "_costvol.bin")
