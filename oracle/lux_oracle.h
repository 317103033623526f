/*
 * lux_oracle.h -- CPU restatement of Lux.Engine's dense-stereo SGM layer, used ONLY as test
 * infrastructure: the parity checker for the HIP library and the timed CPU baseline of bench.py.
 * Nothing in the product library links, loads or calls this code.
 *
 * Reference files restated (paths relative to the mprovenz/lux checkout):
 *   Lux.Engine/Pipeline/Registration/DenseStereoSgm.cs   DenseLayer (Init ranges, line buffers, Compute, Pixel, Rd, WTA)
 *   Lux.Engine/Pipeline/Registration/DenseStereo.cs      BoxFilter3 (:93-159), Guidance (:162-175), MatchingCost and
 *                                                        its scalar form (:217-373), Normalise (:376-381)
 *   Lux.Engine/Pipeline/Registration/DenseStereoSgm.cs   also SkipMask (:64-87), RangeFromPrevious (:91-140),
 *                                                        the layer's depth (:225-231)
 *   Lux.Engine/Pipeline/Registration/DenseUpsampleLayer.cs  ImageShift, NearestResize, BilateralUpsampleFromCollapse
 *   Lux.Engine/Pipeline/Registration/IntelApprox.cs      rcpss, rsqrtss (through the tables read from the instructions)
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
/* The same layer with its visits in a caller-given order: visits[3k..3k+2] = (x, y, cur half), the first n_plus
 * in pass +1, the rest in pass -1. In the reference's order (lo_layer_compute's loop) it gives the same results;
 * the tests use it to run orders a parallel schedule allows. Returns -2 on a visit outside the image. */
int lo_layer_compute_visits(const lo_layer_params* p, const uint8_t* raw_packed, uint16_t* agg_packed,
                            uint16_t* agg_pass1_packed, uint16_t* wta, const int32_t* visits, int32_t n_plus,
                            int32_t n_minus);

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

/* ---- DenseUpsampleLayer (Lux.Engine/Pipeline/Registration/DenseUpsampleLayer.cs) ---- */
/* rcpss / rcpps as an Intel CPU computes them (IntelApprox.cs:21-29): a 4096-entry table on the top 12 mantissa
 * bits scaled by the input's power of two; +-0 and denormals -> +-inf, +-inf -> +-0, NaN -> quiet NaN, a denormal
 * result -> +-0. */
uint32_t lo_rcp_bits(uint32_t b);
/* The local CPU's own rcpss (for pinning lo_rcp_bits on an Intel host). */
uint32_t lo_rcpss_hw(uint32_t b);
/* IntelApprox.RsqrtBits (IntelApprox.cs:32-44) and the local CPU's own rsqrtss (for pinning it on an Intel host). */
uint32_t lo_rsqrt_bits(uint32_t b);
uint32_t lo_rsqrtss_hw(uint32_t b);
/* Test helper: how many bit patterns b in [lo, hi) the model (which 0: rcp, 1: rsqrt) maps differently from the local
 * CPU's instruction. */
uint64_t lo_approx_mismatches(int which, uint64_t lo, uint64_t hi);
/* Homography.Sqrt (Homography.cs:18). */
float lo_homography_sqrt(float x);
/* DenseStereo.Planes (DenseStereo.cs:41-88); calibs: ncal records of K[9], T[3], R[9] floats. Returns 0, or -1 near >=
 * far, -2 a non-positive end, -3 baseline too small, -4 bad arguments; *n_out = the plane count. */
int lo_planes(float near_z, float far_z, const float* calibs, int ncal, float density, int align, float* planes, int cap,
              int* n_out);
/* Mat4D.FlowMatrix (ViewTransform.cs:120-133): column-major float out[16] of (K_b E_b)(K_a E_a)^-1. */
void lo_flow_matrix(const float* a, const float* b, float* out);
/* ImageShift<2, vec4x8ui>(src, {sx, sy}) (:64-104): RGBA8 W x H, separable 4-tap cubic, clamp-to-edge, RNE, saturate. */
int lo_image_shift(const uint8_t* src, int W, int H, float sx, float sy, uint8_t* dst);
/* NearestResize (:108-128): 16.16 steps, RGBA8 sw x sh -> W x H. */
int lo_nearest_resize(const uint8_t* src, int sw, int sh, int W, int H, uint8_t* dst);
/* BilateralUpsampleFromCollapse<2, float, vec4x8ui> (:156-210): lo lw x lh -> out (guide's W x H). */
int lo_bilateral_upsample(const float* lo, int lw, int lh, const uint8_t* guide, int W, int H, const uint8_t* g2, float sigma,
                          float* out);
/* The whole step (:38-52): shift the guide by (0.5, 0.5), resize to lw x lh, invert the depth, upsample, invert.
 * Returns -3 when the guide is not (setW*2) x (setH*2) ("Incorrect output depth size!"). */
int lo_upsample_layer(const float* prev_depth, int pw, int ph, const uint8_t* guide, int W, int H, int setW, int setH,
                      float sigma, float* depth);

/* ---- Plane-sweep matching cost (Lux.Engine/Pipeline/Registration/DenseStereo.cs) ---- */
/* One other camera of the cost context (BuildContext :190-200): its box-filtered RGBA8 image (dense rows), the
 * reference->camera flow matrix (column-major floats, Warp :186) and the channel weights Wq. */
typedef struct {
    const uint8_t* rgba;
    int w, h;
    float warp[16];
    uint16_t wq[4];
} lo_view;
typedef struct {
    const uint8_t* ref;   /* filtered[0], dense RGBA8 rw x rh */
    int rw, rh;
    const lo_view* views;
    int nv;
    const float* planes;  /* metric plane depths, far .. near */
    int n;
    uint8_t caps[4];      /* per-channel caps of the absolute differences ({2, 6, 6, 0}) */
} lo_sweep;
/* MatchingCost (:217-292, its scalar form :295-373, equal to it) for reference image pixel (x, y), planes
 * [start, start+count): u16 costs, wrapping sum over the cameras. A first-plane projection that is NaN is clamped
 * like any other (maxNum/minNum); the reference raises there. Returns -1 on bad arguments. */
int lo_matching_cost(const lo_sweep* s, int x, int y, int start, int count, uint16_t* cost);
/* Normalise (:376-381): raw[d] = (byte)(int)(cost[d] * (0x3d17b426 / nv)), planes [count, cap) = 255. */
void lo_normalise(const uint16_t* cost, int count, int cap, int nv, uint8_t* raw);
/* The raw volume of a layer (DenseLayer.Compute :186-205): layer pixel (x, y) reads image pixel
 * (min(scale*x + scale/2, rw-1), min(scale*y + scale/2, rh-1)); raw is packed pixel by pixel, roundup(count, align)
 * bytes each; skipped pixels (skip != 0) get 0 over their planes and 255 over the padding, as Pixel uses them (:242). */
int lo_layer_sweep_cost(const lo_sweep* s, int W, int H, int scale, int align, const uint16_t* start, const uint16_t* count,
                        const uint8_t* skip, uint8_t* raw, int threads);
/* Guidance (:162-175): the reference image itself when it is w x h, else the pixel at
 * (min(scale*x + scale/2, rw-1), min(scale*y + scale/2, rh-1)). */
int lo_guidance(const uint8_t* ref, int rw, int rh, int w, int h, int scale, uint8_t* out);
/* RangeFromPrevious (DenseStereoSgm.cs:91-140): the plane range of every pixel of a w x h layer from the previous
 * layer's depth (pw x ph) and skip mask. Writes start/count and returns MinLo / MaxHi (MaxHi not rounded). */
int lo_range_from_previous(const float* planes, int n, const float* prev_depth, const uint8_t* prev_skip, int pw, int ph,
                           int w, int h, int margin, uint16_t* start, uint16_t* count, int* min_lo, int* max_hi);
/* The layer's depth (Compute :225-231): Planes[start + first minimum of agg over count]. agg packed as raw. */
int lo_layer_depth(const float* planes, int n, int W, int H, int align, const uint16_t* start, const uint16_t* count,
                   const uint16_t* agg, float* depth);

/* DenseLayer.SkipMask (DenseStereoSgm.cs:64-87): mode 0 all computed; mode 2 one computed pixel per 2x2 cell of
 * each 64 x 64 tile, drawn from std::mt19937 seeded 5489 per tile (the last tile of a row / column takes the
 * remainder). 0 computed, 0xff skipped. Returns -2 for another mode. */
int lo_skip_mask(int mode, int w, int h, uint8_t* out);
/* The n-th output of std::mt19937 seeded with `seed` (the textbook generator). */
uint32_t lo_mt19937(uint32_t seed, int n);

#ifdef __cplusplus
}
#endif
#endif
