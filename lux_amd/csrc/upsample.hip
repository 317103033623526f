// DenseUpsampleLayer on the device (Lux.Engine/Pipeline/Registration/DenseUpsampleLayer.cs:38-52): the finest layer's
// depth is carried to the guide's resolution by a bilateral upsample of its inverse. Every float operation is one
// rounded operation (__fmul_rn and friends, no contraction), in the reference's order, so the result is the
// reference's bit for bit; the reciprocals are Intel's rcpss (the table read from the instruction, rcp_table.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/lux_sgm.h"
#include "engine.h"
#include "rcp_dev.h"

namespace lux {
namespace {

struct Taps {
    float kx[4], ky[4];
    int ix, iy;
};

// g2 = NearestResize(ImageShift<2, vec4x8ui>(guide, {0.5, 0.5}), pw, ph) (:70-128) in one pass: the resize keeps
// the shifted image only at (x * stepX >> 16, y * stepY >> 16), so the shift is evaluated at those pixels alone.
// One output byte per thread: the four source columns' vertical sums over the four clamped rows (acc = v * k + acc
// from the first row), then ((v3 kx3 + v2 kx2) + (v1 kx1 + v0 kx0)), rounded to nearest even and saturated.
__global__ void shifted_resize_kernel(const uint8_t* src, int W, int H, Taps t, int pw, int step_x, int step_y, uint8_t* dst)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= pw * 4) return;
    const int ch = x & 3, px = ((x >> 2) * step_x) >> 16, py = (y * step_y) >> 16;
    int rows[4];
#pragma unroll
    for (int i = 0; i < 4; i++) rows[i] = min(max(py + t.iy - 1 + i, 0), H - 1);
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int cx = min(max(px + t.ix - 1 + k, 0), W - 1);
        float acc = __fmul_rn((float)src[((size_t)rows[0] * W + cx) * 4 + ch], t.ky[0]);
#pragma unroll
        for (int i = 1; i < 4; i++) acc = __fadd_rn(__fmul_rn((float)src[((size_t)rows[i] * W + cx) * 4 + ch], t.ky[i]), acc);
        v[k] = acc;
    }
    const float s = __fadd_rn(__fadd_rn(__fmul_rn(v[3], t.kx[3]), __fmul_rn(v[2], t.kx[2])),
                              __fadd_rn(__fmul_rn(v[1], t.kx[1]), __fmul_rn(v[0], t.kx[0])));
    const int q = __float2int_rn(s);
    dst[(size_t)y * pw * 4 + x] = (uint8_t)min(max(q, 0), 255);
}

// The bilateral weight's fast 2^t (:144-152): n = (int)t + (sign bit ? -1 : 0).
__device__ __forceinline__ float exp2_fast_signbit(float t)
{
    t = t < 128.0f ? t : 128.0f;
    t = t > -126.0f ? t : -126.0f;
    const int n = (int)t + (__float_as_int(t) >> 31);
    const float f = __fsub_rn(t, (float)n);
    const float c1 = __int_as_float(0x3d9fcb52), c2 = __int_as_float(0x3e677e26), c3 = __int_as_float(0x3f322226),
                c4 = __int_as_float(0x3f7ffb19);
    const float p = __fadd_rn(__fmul_rn(__fadd_rn(__fmul_rn(__fadd_rn(__fmul_rn(f, c1), c2), f), c3), f), c4);
    return __int_as_float((int)((uint32_t)__float_as_int(p) + ((uint32_t)n << 23)));
}

__device__ __forceinline__ float spatial_w(float f, int i)
{
    float d = fabsf(__fsub_rn(f, (float)(i * 2)));
    d = d < 1.0f ? d : 1.0f;
    return (int)d == 0 ? 1.0f : __int_as_float(0x3eaaaaab);
}

__device__ __forceinline__ float range_w(float hr, float hg, float hb, uint32_t q, float k, float wyx)
{
    const float d0 = __fsub_rn(hr, (float)(q & 255u)), d1 = __fsub_rn(hg, (float)((q >> 8) & 255u)),
                d2 = __fsub_rn(hb, (float)((q >> 16) & 255u));
    const float s = __fadd_rn(__fmul_rn(d1, d1), __fadd_rn(__fmul_rn(d2, d2), __fmul_rn(d0, d0)));
    return __fmul_rn(__fadd_rn(exp2_fast_signbit(-__fmul_rn(s, k)), __int_as_float(0x322bcc77)), wyx);
}

// InverseDepth (:133-140) of the previous depth, once per low-res pixel.
__global__ void inverse_kernel(const float* src, float* dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = rcp_intel(src[i]);
}

// BilateralUpsampleFromCollapse<2, float, vec4x8ui> (:156-210) of the inverse depth, one output pixel per thread,
// then InverseDepth again. With `conf` the previous confidence goes through the same weights in the same order
// (:49-50, no inverse). Guide and g2 pixels are read as one 32-bit word each (the buffers are 4-byte aligned).
template <bool CONF>
__global__ void bilateral_kernel(const float* inv, const float* conf, int lw, int lh, const uint32_t* guide, int W, int H,
                                 const uint32_t* g2, float k, float* out, float* conf_out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    const int wm1 = lw - 1, hm1 = lh - 1;
    const int iy0 = (y - 1) / 2;
    const float fy = __fadd_rn((float)y, -0.5f);
    const uint32_t gv = guide[(size_t)y * W + x];
    const float hr = (float)(gv & 255u), hg = (float)((gv >> 8) & 255u), hb = (float)((gv >> 16) & 255u);
    const int ix0 = (x - 1) / 2;
    const float fx = __fadd_rn((float)x, -0.5f);
    const int ibx = min(max(ix0, 0), wm1), ix1 = min(max(ix0 + 1, 0), wm1);
    const float wx0 = spatial_w(fx, ibx), wx1 = spatial_w(fx, ix1);
    float av = 0.0f, as = 0.0f, ac = 0.0f;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int iy = min(max(iy0 + j, 0), hm1);
        const float wy = spatial_w(fy, iy);
        const size_t i0 = (size_t)iy * lw + ibx, i1 = (size_t)iy * lw + ix1;
        const float w0 = range_w(hr, hg, hb, g2[i0], k, __fmul_rn(wy, wx0));
        const float w1 = range_w(hr, hg, hb, g2[i1], k, __fmul_rn(wy, wx1));
        const float v0 = __fmul_rn(inv[i0], w0), v1 = __fmul_rn(inv[i1], w1);
        av = __fadd_rn(v1, __fadd_rn(v0, av));
        as = __fadd_rn(w1, __fadd_rn(w0, as));
        if (CONF) ac = __fadd_rn(__fmul_rn(conf[i1], w1), __fadd_rn(__fmul_rn(conf[i0], w0), ac));
    }
    out[(size_t)y * W + x] = rcp_intel(__fdiv_rn(av, as));
    if (CONF) conf_out[(size_t)y * W + x] = __fdiv_rn(ac, as);
}

// The cubic tap (:58-63), host side: plain float operations (no contraction in this build).
float cubic_tap(float x)
{
#pragma clang fp contract(off)
    uint32_t sb = 0x3e2aaaabu;
    float sixth;
    std::memcpy(&sixth, &sb, 4);
    const float x2 = x * x, x3 = x2 * x;
    if (x < 1.0f) {
        const float a = x3 * 9.0f;
        float b = x2 * -15.0f;
        b = b + 6.0f;
        b = b + a;
        return b * sixth;
    }
    if (x < 2.0f) {
        const float a = x3 * -3.0f, b = x2 * 15.0f;
        float c = x * -24.0f;
        c = c + 12.0f;
        c = c + b;
        c = c + a;
        return c * sixth;
    }
    return 0.0f;
}

Taps taps_for(float sx, float sy)
{
    Taps t;
    t.ix = (int)std::floor(sx);
    t.iy = (int)std::floor(sy);
    const float tx = sx - (float)t.ix, ty = sy - (float)t.iy;
    const float ax[4] = {tx + 1.0f, tx, 1.0f - tx, 2.0f - tx}, ay[4] = {ty + 1.0f, ty, 1.0f - ty, 2.0f - ty};
    for (int i = 0; i < 4; i++) {
        t.kx[i] = cubic_tap(ax[i]);
        t.ky[i] = cubic_tap(ay[i]);
    }
    return t;
}

}  // namespace

int enqueue_upsample(Workspace& ws, const float* d_prev, const float* d_prev_conf, int pw, int ph, const uint8_t* d_guide,
                     int W, int H, float sigma, float* d_depth, float* d_conf, hipStream_t s, std::string* err)
{
    void *g2, *inv;
    int rc;
    if ((rc = ws.get("ups_g2", (size_t)pw * ph * 4, &g2, err)) || (rc = ws.get("ups_inv", (size_t)pw * ph * 4, &inv, err)))
        return rc;
    if (((uintptr_t)d_guide & 3u) != 0) {
        if (err) *err = "the guide must be 4-byte aligned";
        return LUX_SGM_EINVAL;
    }
    const Taps t = taps_for(0.5f, 0.5f);  // ImageShift<2>(guide, {0.5f, 0.5f}) (:41)
    const int step_y = (int)((double)H * 65536.0 / (double)ph), step_x = (int)((double)W * 65536.0 / (double)pw);
    shifted_resize_kernel<<<dim3((pw * 4 + 255) / 256, ph), 256, 0, s>>>(d_guide, W, H, t, pw, step_x, step_y, (uint8_t*)g2);
    const size_t np = (size_t)pw * ph;
    inverse_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(d_prev, (float*)inv, np);
    const float k = 0.5f / (sigma * sigma);
    const dim3 grid((W + 255) / 256, H);
    const uint32_t* gd = (const uint32_t*)d_guide;
    if (d_prev_conf && d_conf)
        bilateral_kernel<true><<<grid, 256, 0, s>>>((const float*)inv, d_prev_conf, pw, ph, gd, W, H, (const uint32_t*)g2, k,
                                                    d_depth, d_conf);
    else
        bilateral_kernel<false><<<grid, 256, 0, s>>>((const float*)inv, nullptr, pw, ph, gd, W, H, (const uint32_t*)g2, k,
                                                     d_depth, nullptr);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_status(e, "upsample", err);
}

}  // namespace lux
