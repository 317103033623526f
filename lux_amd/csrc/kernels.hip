// Host-side launchers of liblux_sgm's kernels (declared in kernels.h), and the constant-P2 instantiation of
// the request's launch sequence. The device code is in kernels_impl.h.
#include "kernels_impl.h"

namespace lux {

extern template hipError_t launch_sgm_t<true>(const Geometry& g, const SgmBuffers& b, const SgmOutputs& o, int p1, int p2,
                                              uint16_t* agg_pass1, hipStream_t s, hipEvent_t* m);

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s)
{
    fill_u32<<<grid_for(n, 256), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

int band_tiles(int W, int H)
{
    // at most one extra (short) band from the split at the seed row
    return ((H + kBandRows - 1) / kBandRows + 1) * ((W + kStrip - 1) / kStrip);
}

Geometry geometry(int W, int H, int D, int align)
{
    Geometry g;
    g.W = W; g.H = H; g.D = D;
    g.cap = ((D + align - 1) / align) * align;
    int dpl = 2;
    while (64 * dpl < g.cap) dpl *= 2;
    g.dpl = dpl;
    const int q = dpl > 4 ? dpl : 4;
    g.cs = ((g.cap + q - 1) / q) * q;
    return g;
}

hipError_t launch_census(const void* img, bool img16, int W, int H, int stride, int radius, uint64_t* codes, hipStream_t s)
{
    dim3 blk(16, 16), grd((W + 15) / 16, (H + 15) / 16);
#define LUX_CENSUS(T, R) census_kernel<T, R><<<grd, blk, 0, s>>>((const T*)img, W, H, stride, codes)
    if (img16) {
        if (radius == 1) LUX_CENSUS(uint16_t, 1); else if (radius == 2) LUX_CENSUS(uint16_t, 2); else LUX_CENSUS(uint16_t, 3);
    } else {
        if (radius == 1) LUX_CENSUS(uint8_t, 1); else if (radius == 2) LUX_CENSUS(uint8_t, 2); else LUX_CENSUS(uint8_t, 3);
    }
#undef LUX_CENSUS
    return hipGetLastError();
}

hipError_t launch_cost_ad(const void* L, const void* R, bool img16, int bits, int stride, int adcap, const Geometry& g,
                          uint8_t* raw, hipStream_t s)
{
    dim3 grd((g.W * (g.cs / 4) + kBlock - 1) / kBlock, g.H);
    const int shift = img16 ? bits - 8 : 0;
    if (img16)
        cost_ad_kernel<uint16_t><<<grd, kBlock, 0, s>>>((const uint16_t*)L, (const uint16_t*)R, g.W, g.H, stride, g.D, g.cs,
                                                        shift, adcap, raw);
    else
        cost_ad_kernel<uint8_t><<<grd, kBlock, 0, s>>>((const uint8_t*)L, (const uint8_t*)R, g.W, g.H, stride, g.D, g.cs,
                                                       shift, adcap, raw);
    return hipGetLastError();
}


hipError_t launch_sgm(const Geometry& g, const SgmBuffers& b, const SgmOutputs& o, int p1, int p2, uint16_t* agg_pass1,
                      hipStream_t s, hipEvent_t* m)
{
    if (b.ap.on) return launch_sgm_t<true>(g, b, o, p1, p2, agg_pass1, s, m);
    return launch_sgm_t<false>(g, b, o, p1, p2, agg_pass1, s, m);
}

hipError_t launch_restride_u16(const uint16_t* src, int src_cs, uint16_t* dst, int dst_cs, int cells, int n, hipStream_t s)
{
    const size_t total = (size_t)cells * dst_cs;
    restride_kernel<uint16_t><<<grid_for(total, 256), 256, 0, s>>>(src, src_cs, dst, dst_cs, (size_t)cells, n, 0);
    return hipGetLastError();
}

hipError_t launch_downsample2(const void* src, bool img16, int W, int H, int stride, void* dst, int W2, int H2, hipStream_t s)
{
    dim3 blk(32, 8), grd((W2 + 31) / 32, (H2 + 7) / 8);
    if (img16) downsample2_kernel<uint16_t><<<grd, blk, 0, s>>>((const uint16_t*)src, W, H, stride, (uint16_t*)dst, W2, H2);
    else downsample2_kernel<uint8_t><<<grd, blk, 0, s>>>((const uint8_t*)src, W, H, stride, (uint8_t*)dst, W2, H2);
    return hipGetLastError();
}

hipError_t launch_ranges_from_coarse(const uint16_t* dc, int Wc, int Hc, int W, int H, int r, int D, uint16_t* vstart,
                                     uint16_t* vcap, uint16_t* vcount, hipStream_t s)
{
    const size_t n = (size_t)W * H;
    ranges_from_coarse_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dc, Wc, Hc, W, H, r, D, vstart, vcap, vcount);
    return hipGetLastError();
}

hipError_t launch_cost_census_ranges(const uint64_t* cl, const uint64_t* cr, int W, int H, int cs, const uint16_t* vstart,
                                     const uint16_t* vcount, uint8_t* raw, hipStream_t s)
{
    const size_t n = (size_t)W * H * cs;
    cost_census_ranges_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(cl, cr, W, H, cs, vstart, vcount, raw);
    return hipGetLastError();
}

// The 3x3 box filter of the matcher's inputs (DenseStereo.BoxFilter3, DenseStereo.cs:93-159), per channel:
// (float)sum over the in-image window * (1 / (cols * rows)), rounded to nearest even, saturated to maxval.
template <typename T>
__global__ void box3_kernel(const T* src, int W, int H, int stride, int ch, int maxval, T* dst)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const int x0 = max(x - 1, 0), x1 = min(x + 1, W - 1), y0 = max(y - 1, 0), y1 = min(y + 1, H - 1);
    const float inv = __fdiv_rn(1.0f, __fmul_rn((float)(x1 - x0 + 1), (float)(y1 - y0 + 1)));
    for (int c = 0; c < ch; c++) {
        int sum = 0;
        for (int yy = y0; yy <= y1; yy++)
            for (int xx = x0; xx <= x1; xx++) sum += src[(size_t)yy * stride * ch + (size_t)xx * ch + c];
        const int r = __float2int_rn(__fmul_rn((float)sum, inv));
        dst[((size_t)y * W + x) * ch + c] = (T)min(max(r, 0), maxval);
    }
}

// The same filter on four interleaved 8-bit channels: one 32-bit load per tap, the four sums from its bytes, and the
// reciprocal of the window's size (9, 6 or 4 pixels away from tiny images) as a constant: the literal quotients are
// the correctly rounded ones the division gives.
__global__ void box3_rgba8_kernel(const uint32_t* src, int W, int H, int stride, int maxval, uint32_t* dst)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const int x0 = max(x - 1, 0), x1 = min(x + 1, W - 1), y0 = max(y - 1, 0), y1 = min(y + 1, H - 1);
    const int n = (x1 - x0 + 1) * (y1 - y0 + 1);
    const float inv = n == 9 ? 1.0f / 9.0f : n == 6 ? 1.0f / 6.0f : n == 4 ? 0.25f : __fdiv_rn(1.0f, (float)n);
    int s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int yy = y0; yy <= y1; yy++)
        for (int xx = x0; xx <= x1; xx++) {
            const uint32_t v = src[(size_t)yy * stride + xx];
            s0 += (int)(v & 0xffu);
            s1 += (int)((v >> 8) & 0xffu);
            s2 += (int)((v >> 16) & 0xffu);
            s3 += (int)(v >> 24);
        }
    const auto q = [&](int sum) { return (uint32_t)min(max(__float2int_rn(__fmul_rn((float)sum, inv)), 0), maxval); };
    dst[(size_t)y * W + x] = q(s0) | (q(s1) << 8) | (q(s2) << 16) | (q(s3) << 24);
}

hipError_t launch_box3(const void* src, bool img16, int W, int H, int stride, int ch, int maxval, void* dst, hipStream_t s)
{
    dim3 blk(32, 8), grd((W + 31) / 32, (H + 7) / 8);
    if (!img16 && ch == 4 && ((uintptr_t)src & 3) == 0 && ((uintptr_t)dst & 3) == 0) {
        box3_rgba8_kernel<<<grd, blk, 0, s>>>((const uint32_t*)src, W, H, stride, maxval, (uint32_t*)dst);
        return hipGetLastError();
    }
    if (img16) box3_kernel<uint16_t><<<grd, blk, 0, s>>>((const uint16_t*)src, W, H, stride, ch, maxval, (uint16_t*)dst);
    else box3_kernel<uint8_t><<<grd, blk, 0, s>>>((const uint8_t*)src, W, H, stride, ch, maxval, (uint8_t*)dst);
    return hipGetLastError();
}

// Skipped pixels (DenseLayer.Skip != 0) aggregate raw 0 over their planes and 255 over the padding (:242).
__global__ void apply_skip_kernel(uint8_t* raw, const uint8_t* skip, const uint16_t* vcount, int count, int cs, size_t cells)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * cs) return;
    const size_t c = i / cs;
    if (!skip[c]) return;
    const int j = (int)(i - c * cs);
    raw[i] = j < (vcount ? (int)vcount[c] : count) ? 0 : 255;
}

hipError_t launch_apply_skip(uint8_t* raw, const uint8_t* skip, const uint16_t* vcount, int count, int cs, size_t cells,
                             hipStream_t s)
{
    const size_t n = cells * (size_t)cs;
    apply_skip_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(raw, skip, vcount, count, cs, cells);
    return hipGetLastError();
}

hipError_t launch_unpack_u8(const uint8_t* src, const uint64_t* off, const uint16_t* cap, uint8_t* dst, int cs, size_t cells,
                            hipStream_t s)
{
    const size_t total = cells * (size_t)cs;
    unpack_kernel<uint8_t><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(src, off, cap, dst, cs, cells, (uint8_t)255);
    return hipGetLastError();
}

hipError_t launch_pack_u16(const uint16_t* src, const uint64_t* off, const uint16_t* cap, uint16_t* dst, int cs, size_t cells,
                           hipStream_t s)
{
    const size_t total = cells * (size_t)cs;
    pack_kernel<uint16_t><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(src, off, cap, dst, cs, cells);
    return hipGetLastError();
}

hipError_t launch_pack_u8(const uint8_t* src, const uint64_t* off, const uint16_t* cap, uint8_t* dst, int cs, size_t cells,
                          hipStream_t s)
{
    const size_t total = cells * (size_t)cs;
    pack_kernel<uint8_t><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(src, off, cap, dst, cs, cells);
    return hipGetLastError();
}

hipError_t launch_restride_u8(const uint8_t* src, int src_cs, uint8_t* dst, int dst_cs, int cells, int n, uint8_t fill,
                              hipStream_t s)
{
    const size_t total = (size_t)cells * dst_cs;
    restride_kernel<uint8_t><<<grid_for(total, 256), 256, 0, s>>>(src, src_cs, dst, dst_cs, (size_t)cells, n, fill);
    return hipGetLastError();
}

}  // namespace lux
