/*
 * lux_oracle.h -- CPU restatement of Lux.Engine's dense-stereo SGM layer, used ONLY as test
 * infrastructure: the parity checker for the HIP library and the timed CPU baseline of bench.py.
 * Nothing in the product library links, loads or calls this code.
 *
 * Reference files restated (paths relative to the mprovenz/lux checkout):
 *   Lux.Engine/Pipeline/Registration/DenseStereoSgm.cs   DenseLayer (Init ranges, line buffers, Compute, Pixel, Rd, WTA)
 *   Lux.Engine/Pipeline/Registration/DenseStereo.cs      BoxFilter3 (:93-159), out-of-range plane rule (:273-274),
 *                                                        padding lanes = 255 (:376-381)
 *   Lux.Engine/Pipeline/Parallax/RealStereo.cs:229-233,418-421  two-direction check pattern (L/R check)
 *   Lux.Engine/Pipeline/DepthRefine/CrossViewCost.cs:197-202    parabola vertex (subpixel)
 * Census and the truncated-AD cost of a rectified pair are project-defined front ends (the reference
 * SGM has no census stage); their definitions are stated at each function.
 */
#ifndef LUX_ORACLE_H
#define LUX_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Census transform: for a (2r+1)x(2r+1) window (r = 2 or 3), bit k (LSB first, raster order of the
 * window with the centre skipped) is 1 when I(neighbour) < I(centre). Neighbour coordinates are clamped
 * to the image (replicated border). img16 != 0 selects uint16 pixels, else uint8. stride in elements. */
int lo_census(const void* img, int img16, int W, int H, int stride, int radius, uint64_t* codes);

/* Census matching cost, cell-major raw[(y*W + x)*cap + d]:
 *   d < D : popcount(cl(x,y) ^ cr(max(x-d, 0), y))   (a plane whose right pixel falls left of the image
 *                                                    repeats the last in-image plane, DenseStereo.cs:273-274)
 *   D <= d < cap : 255                               (padding lanes, DenseStereo.cs:380)
 * Row-parallel over `threads` workers, as DenseStereoSgm.cs:186-205. */
int lo_cost_census(const uint64_t* cl, const uint64_t* cr, int W, int H, int D, int cap, uint8_t* raw, int threads);

/* Truncated 3x3 absolute-difference cost of single-channel images (project definition, the
 * single-channel analogue of DenseStereo.MatchingCost's capped 3x3 SAD):
 *   raw = sum over the 3x3 window (rows/cols clamped) of min(|L - R| >> shift, adcap), shift = bits - 8
 *   for 16-bit input (bits = 10..16) and 0 for 8-bit input; right window centred on max(x-d, 0).
 * adcap <= 28 keeps the sum within a byte (9*28 = 252, DenseStereo.cs:209). Padding lanes 255. */
int lo_cost_ad(const void* L, const void* R, int img16, int bits, int W, int H, int stride, int D, int cap,
               int adcap, uint8_t* raw, int threads);

/* BoxFilter3 (DenseStereo.cs:93-159) on a single-channel image: per pixel float(sum of the in-image
 * 3x3 window) * (1/(cols*rows)) rounded to nearest even, saturated to [0, maxval]. */
int lo_box3(const void* src, int img16, int W, int H, int maxval, void* dst);

/* DenseLayer restatement (DenseStereoSgm.cs:13-325). Per-pixel plane ranges start/count; cap is
 * roundup(count, align). Raw and Agg are packed pixel by pixel, cap entries each, in raster order
 * (raw entries of skipped pixels are ignored). P2: p2_const >= 0 uses that integer for every path step
 * (the reference with zero guidance weights gives trunc(P1*P2Scale*0x3f7ffb19)); p2_const < 0 computes
 * the image-adaptive P2 of Pixel (:265-274) from the RGBA8 guidance image, gw and p2_scale. */
typedef struct {
    int W, H, T, align, xsw;       /* xsw: LUX_SGM_XSW=1 behaviour (pass +1 x-start clamp at W) */
    int p1;                        /* P1 (u16) */
    int p2_const;                  /* >= 0: constant P2; < 0: adaptive */
    float p2_scale;                /* P2Scale (adaptive mode) */
    float gw[4];                   /* guidance weights (adaptive mode) */
    const uint8_t* guidance;       /* RGBA8 W*H*4, may be NULL in constant mode */
    const uint16_t* start;         /* W*H */
    const uint16_t* count;         /* W*H */
    const uint8_t* skip;           /* W*H, 0 = computed, else skipped; may be NULL */
} lo_layer_params;

/* Returns 0 on success, -1 on bad arguments, -2 when the reference would fault (out-of-range write). */
int lo_layer_compute(const lo_layer_params* p, const uint8_t* raw_packed, uint16_t* agg_packed,
                     uint16_t* agg_pass1_packed /* may be NULL */, uint16_t* wta /* W*H plane index */);

/* The image-adaptive P2 of one path step (DenseStereoSgm.cs:267-273): S = (|d0|+|d2|)+|d1| with
 * d = (g - q)*gw, e = clamp(-S, -126, 128), E = fast exp2(e) (cubic + exponent add), trunc(p2Base*E). */
int lo_p2_adaptive(float p2Base, const float g[4], const float q[4], const float gw[4]);

/* Test hook: 1 makes every path step take the reference's scalar form (:300-313), 0 (default) lets
 * steps that meet the vector form's conditions (:280-282) take the vector form. */
void lo_set_force_scalar(int v);

/* Winner-take-all + subpixel + left/right check over a uniform-range aggregate volume
 * agg[(y*W + x)*cap + d], planes [0, D):
 *   disp(x)  = first minimum of agg over d in [0, D) (strict <, DenseStereoSgm.cs:225-231)
 *   sub(x)   = disp + parabola offset 0.5*(l-r)/(l-2c+r) in float32 when 0 < disp < D-1 and the
 *              denominator is > 0, else disp (CrossViewCost.cs:197-202)
 *   dR(xr)   = first minimum over d in [0, D), xr+d < W, of agg(xr+d, d)
 *   valid(x) = x - disp >= 0 && |disp - dR(x - disp)| <= tol (RealStereo.cs:418-421) */
int lo_wta_sub_lr(const uint16_t* agg, int W, int H, int D, int cap, int tol, uint16_t* disp, float* sub,
                  uint8_t* valid);

/* Whole rectified pipeline on the CPU (census or AD cost -> layer -> WTA/subpixel/LR), the timed CPU
 * baseline. cost: 0 census, 1 AD. Returns 0 on success. agg may be NULL. times_ms[3] (may be NULL):
 * cost, aggregation, WTA. */
int lo_rectified(const void* left, const void* right, int img16, int bits, int W, int H, int stride,
                 int cost, int radius, int adcap, int D, int align, int p1, int p2, int T, int xsw, int tol,
                 int threads, uint16_t* disp, float* sub, uint8_t* valid, uint16_t* agg, double* times_ms);

#ifdef __cplusplus
}
#endif
#endif
