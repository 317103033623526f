/*
 * lux_sgm.h -- C ABI of the MI355X (gfx950) dense-stereo SGM library (liblux_sgm.so).
 *
 * Drop-in for the depth hot path of Lux.Engine (mprovenz/lux): the body of DenseLayer.Compute
 * (Lux.Engine/Pipeline/Registration/DenseStereoSgm.cs:181-232) and a rectified census/AD front end that
 * feeds it. Plain C types only; no pointer is kept after a call returns. Status codes: 0 = OK, < 0 = error
 * class (the message is available from lux_sgm_last_error()). A C# caller binds it over P/Invoke, as
 * Lux.Engine/Mtp/GPhoto2/GPhoto2Interop.cs:9-51 and Lux.DepthModel/OnnxDepthNetwork.cs:46-95 bind native code.
 *
 * Threading: contexts are independent; one context is used by one thread at a time (calls on one context
 * are serialised internally). Results do not depend on concurrency (CaptureSession.cs:108-109).
 */
#ifndef LUX_SGM_H
#define LUX_SGM_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Error classes. LUX_SGM_EUNSUPPORTED mirrors the reference's NotSupportedException
 * (DenseStereoSgm.cs:69, :343). */
#define LUX_SGM_OK 0
#define LUX_SGM_EINVAL (-1)
#define LUX_SGM_EUNSUPPORTED (-2)
#define LUX_SGM_ENOMEM (-3)
#define LUX_SGM_EDEVICE (-4)
#define LUX_SGM_ECANCELED (-5)

#define LUX_SGM_COST_CENSUS 0
#define LUX_SGM_COST_AD 1

typedef struct lux_sgm_ctx lux_sgm_ctx;

/* Parameters of one rectified pair. P1/P2 are the integer penalties of every path step; the reference's
 * image-adaptive P2 reduces to trunc(P1*P2Scale*0x3f7ffb19) with zero guidance weights
 * (DenseStereoSgm.cs:267-274), so P2 here is that effective integer. `tile` is the reference's
 * wavefront tile T (DenseStereo.cs:26; even values only), whose border side effects
 * (DenseStereoSgm.cs:206-222) are reproduced bit for bit. */
typedef struct {
    int32_t width, height;     /* pixels */
    int32_t stride;            /* elements per image row, >= width */
    int32_t bits;              /* 8: uint8 pixels; 9..16: uint16 pixels with that many significant bits */
    int32_t cost;              /* LUX_SGM_COST_CENSUS or LUX_SGM_COST_AD */
    int32_t census_radius;     /* 1, 2 (5x5) or 3 (7x7) */
    int32_t ad_cap;            /* AD: per-pixel cap, <= 28 */
    int32_t disparities;       /* D: planes 0..D-1 (DenseLayer Count) */
    int32_t align;             /* lanes per block (DenseLayer Align): cap = roundup(D, align) */
    int32_t p1, p2;            /* penalties, 0..65535 */
    int32_t tile;              /* T of the reference schedule, even, >= 2 */
    int32_t xs_clamp_w;        /* 1: LUX_SGM_XSW=1 behaviour (DenseStereoSgm.cs:31,217) */
    int32_t lr_tolerance;      /* |dL - dR| <= tol is valid (RealStereo.cs:418-421) */
} lux_sgm_rect_params;

/* Library version (major*10000 + minor*100 + patch). */
int lux_sgm_version(void);

/* Create a context on HIP device `device`. hbm_budget bounds the bytes the context may hold on the device
 * (0 = no bound); a request above it fails with LUX_SGM_ENOMEM, as the reference's 1 GiB rule refuses a
 * layer (DenseStereoSgm.cs:330,339-344). Replaces constructing DenseLayer/DenseStereoPyramid state. */
int lux_sgm_create(int device, uint64_t hbm_budget, lux_sgm_ctx** out);
int lux_sgm_destroy(lux_sgm_ctx* ctx);

/* UTF-8 message of the last failed call on ctx (owned by ctx, valid until the next call on it). With
 * ctx == NULL: the last failure of lux_sgm_create on this thread. */
const char* lux_sgm_last_error(const lux_sgm_ctx* ctx);

/* Device bytes a rectified pair needs (the DenseLayer.MemoryEstimate analogue,
 * DenseStereoSgm.cs:142-156). */
int lux_sgm_rect_bytes(const lux_sgm_rect_params* p, uint64_t* bytes);

/* Rectified pair from host memory: cost volume (census or AD) -> 8-path SGM with the reference's
 * schedule -> first-minimum WTA + parabolic subpixel + left/right check.
 *   left/right: uint8 or uint16 images (p->bits), p->stride elements per row.
 *   disp   [H*W] uint16  integer disparity (DenseStereoSgm.cs:225-231)   (may be NULL)
 *   sub    [H*W] float   disparity + parabola offset                     (may be NULL)
 *   valid  [H*W] uint8   1 where the L/R check passes                     (may be NULL)
 *   agg    [H*W*cap] uint16 final aggregate, cap = roundup(D, align)      (may be NULL)
 * Host buffers are read/written only during the call. */
int lux_sgm_rectified(lux_sgm_ctx* ctx, const lux_sgm_rect_params* p, const void* left, const void* right,
                      uint16_t* disp, float* sub, uint8_t* valid, uint16_t* agg);

/* Same, with every pointer in device memory and work enqueued on `stream` (a hipStream_t, NULL = the
 * context's stream). Returns once the work is enqueued; the caller synchronises the stream. agg (if not
 * NULL) receives the aggregate with cell stride cap. */
int lux_sgm_rectified_device(lux_sgm_ctx* ctx, const lux_sgm_rect_params* p, const void* d_left, const void* d_right,
                             uint16_t* d_disp, float* d_sub, uint8_t* d_valid, uint16_t* d_agg, void* stream);

/* lux_sgm_rectified_device without device allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory during the call (code under HIP graph capture, serving loops that forbid hipMalloc). The
 * request's buffers (cost volume, aggregate, census codes, line states, border visit lists) come from one static
 * device array of 96 MiB per device, held by the library from the first context on that device until the process
 * ends; a request that needs more (a 640x480 D=64 census request takes about 68 MiB) fails with LUX_SGM_ENOMEM and
 * nothing enqueued. The call allocates nothing on the device, creates no stream, and runs the border chains in order
 * on `stream`. Every call on a device shares that device's array, so calls that overlap on one device race: serialise
 * them (one stream, or one call at a time). Results are those of lux_sgm_rectified_device, bit for bit. */
int lux_sgm_rectified_device_noalloc(lux_sgm_ctx* ctx, const lux_sgm_rect_params* p, const void* d_left,
                                     const void* d_right, uint16_t* d_disp, float* d_sub, uint8_t* d_valid,
                                     uint16_t* d_agg, void* stream);

/* Parameters of one DenseLayer with a uniform plane range (start 0, count = planes, cap =
 * roundup(planes, align) for every pixel: the layer-0 form of DenseLayer.Init, DenseStereoSgm.cs:43-47). */
typedef struct {
    int32_t width, height;
    int32_t planes;            /* Count */
    int32_t align;             /* Align (8 or 2) */
    int32_t p1, p2;            /* P1 and the effective constant P2 */
    int32_t tile;              /* P.Tile, even */
    int32_t xs_clamp_w;        /* LUX_SGM_XSW */
} lux_sgm_layer_params;

/* DenseLayer.Compute (DenseStereoSgm.cs:181-232) over a precomputed raw cost volume (the bytes
 * DenseLayer.Raw holds, as after the Threads > 1 row-parallel stage, :186-205), host memory:
 *   raw       [H*W*cap] uint8  (padding lanes [planes, cap) = 255, DenseStereo.cs:380)
 *   agg       [H*W*cap] uint16 out: DenseLayer.Agg                       (may be NULL)
 *   agg_pass1 [H*W*cap] uint16 out: DenseLayer.AggPass1 (KeepPass1)      (may be NULL)
 *   wta       [H*W]     uint16 out: plane index start+arg of the WTA      (may be NULL)
 * Skipped pixels (SkipMask, :65-87) need nothing here: their raw bytes (0 in range, 255 padding, :242) are
 * part of `raw`. Per-pixel ranges: lux_sgm_compute_layer_ranges. */
int lux_sgm_compute_layer(lux_sgm_ctx* ctx, const lux_sgm_layer_params* p, const uint8_t* raw, uint16_t* agg,
                          uint16_t* agg_pass1, uint16_t* wta);

/* DenseLayer.Compute with per-pixel plane ranges (DenseLayer Start/Count/Cap, DenseStereoSgm.cs:15-26, :43-58, as
 * RangeFromPrevious leaves them, :91-140). start/count: [H*W] uint16; cap_i = roundup(count_i, align).
 *   raw       jagged: cap_i bytes per pixel in raster order (DenseLayer.Raw[i] concatenated)
 *   agg       out, jagged like raw, uint16 (DenseLayer.Agg[i])                       (may be NULL)
 *   agg_pass1 out, the same after pass +1 (DenseLayer.AggPass1)                      (may be NULL)
 *   wta       [H*W] out: start + first-minimum index over [0, count) (:225-231)       (may be NULL)
 * p->planes is the layer's plane count (start + count <= planes). The path sweeps keep each path vector in
 * its cell's own plane frame and realign it to the next cell's start; that is the reference's line-buffer
 * result as long as no range's padding reaches the last word of its path segment: start - MinLo + cap <=
 * MaxHi for every pixel (MinLo = min start, MaxHi = roundup(max(start + count), align)). A layer that breaks
 * it (the reference then writes into the next segment and reads those words back), has a cap above 512 or an
 * odd tile runs on the line-buffer-literal path, which holds the reference's line buffers word for word; caps
 * above 1024 get LUX_SGM_EUNSUPPORTED. */
int lux_sgm_compute_layer_ranges(lux_sgm_ctx* ctx, const lux_sgm_layer_params* p, const uint16_t* start,
                                 const uint16_t* count, const uint8_t* raw, uint16_t* agg, uint16_t* agg_pass1,
                                 uint16_t* wta);

/* The 3x3 box filter the pyramid applies to its inputs (DenseStereo.BoxFilter3, DenseStereo.cs:93-159): per
 * channel, the in-image window's sum as float times 1/(cols*rows), rounded to nearest even and saturated to the
 * pixel range. channels 1 (uint8 or uint16 with `bits`) or 4 (RGBA8); stride in pixels; dst is dense
 * (width*channels per row). _device: device pointers, enqueued on `stream` (NULL = the context's). */
int lux_sgm_box3_device(lux_sgm_ctx* ctx, const void* d_src, int32_t bits, int32_t channels, int32_t width, int32_t height,
                        int32_t stride, void* d_dst, void* stream);
int lux_sgm_box3(lux_sgm_ctx* ctx, const void* src, int32_t bits, int32_t channels, int32_t width, int32_t height,
                 int32_t stride, void* dst);

/* ---- Plane-sweep matching cost of the L16 dense stereo (DenseStereo.cs:17-381) ---- */
#define LUX_SGM_MAX_VIEWS 16
/* One other camera of the cost context (DenseStereo.CostContext, BuildContext :190-200). */
typedef struct lux_sgm_view {
    const uint8_t* rgba;   /* box-filtered RGBA8 image (BoxFilter3), `height` dense rows of `width` pixels, >= 4 x 4 */
    int32_t width, height;
    float warp[16];        /* Warp(reference, camera) = Mat4D.FlowMatrix, column-major (:186) */
    uint16_t wq[4];        /* channel weights: {8160, 680, 680, 0} colour, {12240, 0, 0, 0} gray (:197) */
} lux_sgm_view;
typedef struct lux_sgm_sweep {
    const uint8_t* ref;    /* the box-filtered reference image, dense RGBA8 */
    int32_t ref_width, ref_height;
    int32_t n_views;       /* 1 .. LUX_SGM_MAX_VIEWS */
    const lux_sgm_view* views;
    const float* planes;   /* DenseStereo.Planes: metric depths, far .. near */
    int32_t n_planes;
    uint8_t caps[4];       /* StereoParams.Caps, {2, 6, 6, 0} */
} lux_sgm_sweep;

/* The raw cost volume of one layer as DenseLayer.Compute produces it before the sweeps (DenseStereoSgm.cs:186-205):
 * MatchingCost + Normalise (DenseStereo.cs:217-381) for every computed pixel. Layer pixel (x, y) reads reference pixel
 * (min(scale*x + scale/2, ref_width-1), min(scale*y + scale/2, ref_height-1)). start/count: the per-pixel plane ranges
 * (both null: every pixel searches all planes); skip: 0 computed, else skipped (may be null). raw: packed pixel by
 * pixel, roundup(count, align) bytes each, raster order; skipped pixels get 0 over their planes and 255 over the
 * padding (the bytes Pixel uses for them, :242). Bit-exact with the reference, the vector and scalar forms alike. */
int lux_sgm_sweep_cost(lux_sgm_ctx* ctx, const lux_sgm_sweep* sweep, int32_t width, int32_t height, int32_t scale,
                       int32_t align, const uint16_t* start, const uint16_t* count, const uint8_t* skip, uint8_t* raw);

/* RangeFromPrevious (DenseStereoSgm.cs:91-140): the plane range of every pixel of a width x height layer from the
 * previous layer's depth [ph*pw] and skip mask (may be null): the plane index of each previous pixel (rcpss against
 * the midpoints of neighbouring planes), its min / max over the 4 x 4 window of computed pixels around the resampled
 * position, widened by `margin` below and one plane above. Writes start / count (16 bits, as the reference stores
 * them) and MinLo / MaxHi (MaxHi not rounded to the alignment). */
int lux_sgm_range_from_previous(lux_sgm_ctx* ctx, const float* planes, int32_t n, const float* prev_depth,
                                const uint8_t* prev_skip, int32_t pw, int32_t ph, int32_t width, int32_t height,
                                int32_t margin, uint16_t* start, uint16_t* count, int32_t* min_lo, int32_t* max_hi);

/* One layer of the L16 dense stereo: StereoParams of the layer plus its size. */
typedef struct lux_sgm_dense_params {
    int32_t width, height;  /* (ref_width + scale - 1) / scale, (ref_height + scale - 1) / scale */
    int32_t scale;          /* StereoParams.Scale: 32, 16, 8, 4, 2, 1 */
    int32_t tile;           /* StereoParams.Tile: 4, 8, 16, 32, 64, 64 */
    int32_t align;          /* 8, or 2 in the pyramid's low-memory mode */
    int32_t p1;             /* StereoParams.P1 (1) */
    float p2_scale;         /* StereoParams.P2Scale (500) */
    float guidance[4];      /* StereoParams.Guidance: the adaptive-P2 weights */
} lux_sgm_dense_params;

/* DenseLayer.Compute (DenseStereoSgm.cs:181-232) of a layer whose ranges and skip mask are set (Init, :34-62):
 * the plane-sweep cost of every computed pixel, the image-adaptive eight-path aggregation (the guidance image is the
 * reference subsampled at the layer's scale, :39), the first-minimum WTA and depth = Planes[start + arg] [H*W].
 * start/count both null: every pixel searches all planes (layer 0). skip may be null. agg / agg_pass1 (packed like
 * the raw volume) and wta (the plane index) may be null. Layers whose padding passes the line-buffer segment, with
 * caps of 513..1024 lanes or an odd tile run on the line-buffer-literal path (lux_sgm_compute_layer_ranges).
 * Returns LUX_SGM_EUNSUPPORTED for ranges above 1024 lanes and for more than 1024 planes at layer 0. */
int lux_sgm_dense_layer(lux_sgm_ctx* ctx, const lux_sgm_sweep* sweep, const lux_sgm_dense_params* params,
                        const uint16_t* start, const uint16_t* count, const uint8_t* skip, float* depth, uint16_t* agg,
                        uint16_t* agg_pass1, uint16_t* wta);

/* DenseStereoPyramid.Run (DenseStereoSgm.cs:328-355) on the device: the images are box-filtered once, then each
 * level is a DenseLayer Init (ranges from the previous level's depth and skip mask, the skip mask of its mode) and
 * Compute. From level 1 on, the memory estimate (DenseLayer.MemoryEstimate, :142-156) picks align 8 if it stays under
 * `budget`, else align 2 with the align-2 planes; above the budget with both it fails (LUX_SGM_EUNSUPPORTED, the
 * reference's NotSupportedException). */
typedef struct lux_sgm_dense_pyramid_params {
    int32_t levels;         /* 6 */
    int32_t scale[8];       /* 32, 16, 8, 4, 2, 1 */
    int32_t tile[8];        /* 4, 8, 16, 32, 64, 64 */
    int32_t skip_mode[8];   /* 0, 0, 0, 0, 2, 2 */
    int32_t p1;             /* 1 */
    float p2_scale;         /* 500 */
    float guidance[4];      /* c/6, c/16, c/16, 0 with c = 0x3ef6384f */
    int32_t margin;         /* 1 */
    uint8_t caps[4];        /* 2, 6, 6, 0 */
    uint64_t budget;        /* 1 GiB */
} lux_sgm_dense_pyramid_params;
/* StereoParams.L16Pyramid() and DenseStereoPyramid.Budget. */
void lux_sgm_dense_pyramid_defaults(lux_sgm_dense_pyramid_params* params);
/* images[0] is the reference; every image is RGBA8 with dense rows (not yet filtered). gray[k] selects the gray
 * weights for camera k (may be null: colour). warps: (n_images - 1) x 16 floats, Warp(reference, camera k) for
 * k = 1.. planes8 / planes2: DenseStereo.Planes(near, far, calibs, density, 8 / 2) (planes2 may be null when the
 * budget is never exceeded). depths[i] (may be null): level i's depth, ((W0 + s - 1) / s) x ((H0 + s - 1) / s) for its
 * scale s. aligns[i] (may be null) receives each level's alignment. */
int lux_sgm_dense_pyramid(lux_sgm_ctx* ctx, const lux_sgm_dense_pyramid_params* params, int32_t n_images,
                          const uint8_t* const* images, const int32_t* widths, const int32_t* heights, const uint8_t* gray,
                          const float* warps, const float* planes8, int32_t n8, const float* planes2, int32_t n2,
                          float* const* depths, int32_t* aligns);
/* The same with the images and depths in device memory (every other argument stays on the host); the images are
 * read, not changed. Synchronises the context's stream before returning. */
int lux_sgm_dense_pyramid_device(lux_sgm_ctx* ctx, const lux_sgm_dense_pyramid_params* params, int32_t n_images,
                                 const uint8_t* const* d_images, const int32_t* widths, const int32_t* heights,
                                 const uint8_t* gray, const float* warps, const float* planes8, int32_t n8,
                                 const float* planes2, int32_t n2, float* const* d_depths, int32_t* aligns);

/* DenseUpsampleLayer.Run (DenseUpsampleLayer.cs:35-48): the finest layer's depth prev_depth [ph*pw] float to the
 * guide's resolution. The RGBA8 guide [height*width*4] is shifted by half a pixel with the 4-tap cubic, resized to
 * pw x ph (16.16 nearest), and steers a bilateral upsample (sigma, 12 in the reference) of the depth's inverse (rcpss);
 * depth [height*width] is the inverse of the result. prev_conf (may be null; then conf is not written) goes through
 * the same bilateral weights with no inverse into conf. width/height must be 2*set_w / 2*set_h, else EINVAL
 * "Incorrect output depth size!". Every float operation is the reference's, in its order, and rcpss is Intel's: the
 * outputs are the reference's bit for bit. _device: device pointers, enqueued on `stream`. */
int lux_sgm_upsample_layer(lux_sgm_ctx* ctx, const float* prev_depth, const float* prev_conf, int32_t pw, int32_t ph,
                           const uint8_t* guide, int32_t width, int32_t height, int32_t set_w, int32_t set_h, float sigma,
                           float* depth, float* conf);
int lux_sgm_upsample_layer_device(lux_sgm_ctx* ctx, const float* d_prev_depth, const float* d_prev_conf, int32_t pw,
                                  int32_t ph, const uint8_t* d_guide, int32_t width, int32_t height, int32_t set_w,
                                  int32_t set_h, float sigma, float* d_depth, float* d_conf, void* stream);

/* ---- Layer set-up arithmetic of the L16 path (host, computed once per capture) ---- */
/* One camera's calibration as Lux.Engine's CalibData holds it (Triangulator.cs:8-19): intrinsics K (3x3, row-major),
 * translation T, rotation R (3x3, row-major), in the order and layout of its 84-byte record (CalibData.FromBytes). */
typedef struct lux_sgm_calib {
    float K[9];
    float T[3];
    float R[9];
} lux_sgm_calib;

/* DenseStereo.Planes (DenseStereo.cs:41-88): the metric plane depths far .. near, uniform in 1/z, their number
 * n = trunc((1/near - 1/far) * density * B * K0[0]) (B the largest distance of a camera centre from camera 0's, through
 * Homography.Sqrt), at most 4096, rounded up to `align`; 1/z from rcpss refined by one Newton step in groups of eight,
 * the remainder by division, the last plane exactly `near`. *n receives n; planes (cap entries, may be NULL for a size
 * query) the depths. LUX_SGM_EINVAL with the reference's InvalidOperationException messages ("The near end has to be
 * smaller than the fard end!", "Both the near and far end have to be positive!", "Baseline too small, check
 * calibration."), readable from lux_sgm_calib_last_error(). The result is the reference's bit for bit. */
int lux_sgm_planes(float near_z, float far_z, const lux_sgm_calib* calibs, int32_t n_calibs, float density, int32_t align,
                   float* planes, int32_t cap, int32_t* n);
/* Mat4D.FlowMatrix (ViewTransform.cs:120-133), the warp DenseStereo.Warp / BuildContext hand the cost
 * (DenseStereo.cs:180-199): M = (K_b E_b)(K_a E_a)^-1 in double with the reference's operation order (its products and
 * adjugate inverse), rounded to float, column-major out[16]: the `warp` of a lux_sgm_view. */
int lux_sgm_flow_matrix(const lux_sgm_calib* a, const lux_sgm_calib* b, float* out);
/* Homography.Sqrt (Homography.cs:18): x * rsqrtss(x) with one Newton step, 0 for 0. */
float lux_sgm_homography_sqrt(float x);
/* Intel's rcpss and rsqrtss as bit functions (IntelApprox.RcpBits / RsqrtBits, IntelApprox.cs:21-44): the tables of the
 * reference machine and its special cases, the same on every host. */
uint32_t lux_sgm_rcpss_bits(uint32_t bits);
uint32_t lux_sgm_rsqrtss_bits(uint32_t bits);
/* Message of the last failed lux_sgm_planes call on this thread. */
const char* lux_sgm_calib_last_error(void);

/* The skip mask of a layer (DenseLayer.SkipMask, DenseStereoSgm.cs:64-87): out[H*W] = 0 for computed pixels, 0xff
 * for skipped ones. Mode 0: all computed; mode 2: per 64x64 tile, a generator seeded 5489 picks one pixel of every
 * 2x2 cell; other modes: LUX_SGM_EUNSUPPORTED (the reference's NotSupportedException, :69). Host only. */
int lux_sgm_skip_mask(int32_t mode, int32_t w, int32_t h, uint8_t* out);
/* The n-th output (n >= 1) of the 32-bit Mersenne twister seeded with `seed` (Mt19937, FundamentalMatrixFilter.cs:7-30). */
uint32_t lux_sgm_mt19937_draw(uint32_t seed, uint32_t n);

/* Image-adaptive P2 (DenseStereoSgm.cs:267-274): every path step takes P2 = (int)(P1 * p2_scale * E),
 * E = fast exp2 of -S, S = (|d0| + |d2|) + |d1|, d_c = (g_c - q_c) * weights[c], g the RGBA guidance of the cell
 * and q that of the cell whose line-buffer state the step reads (0 where nothing was written). */
typedef struct {
    const uint8_t* guidance;   /* [H*W*4] RGBA8 at layer resolution (DenseLayer.Guidance, DenseStereo.cs:163-177) */
    float weights[4];          /* StereoParams.Guidance (DenseStereo.cs:17-21) */
    float p2_scale;            /* StereoParams.P2Scale */
} lux_sgm_guidance;

/* DenseLayer.Compute in full: per-pixel ranges when start/count are given (else the uniform range of p), a
 * skip mask when skip is given ([H*W], non-zero = skipped: such pixels aggregate raw 0 over their count and 255
 * over the padding, :242, whatever their bytes in raw), adaptive P2 when g is given (else the constant p->p2).
 * raw/agg/agg_pass1 are jagged with ranges, [H*W*cap] without; wta as in the two calls above. */
int lux_sgm_compute_layer_ex(lux_sgm_ctx* ctx, const lux_sgm_layer_params* p, const uint16_t* start,
                             const uint16_t* count, const uint8_t* skip, const lux_sgm_guidance* g, const uint8_t* raw,
                             uint16_t* agg, uint16_t* agg_pass1, uint16_t* wta);

/* Coarse-to-fine rectified request: `levels` levels of halved resolution (2x2 mean), the coarsest searching
 * all of its planes, every finer level searching 2*radius planes starting at 2*d - radius (clamped) around
 * twice the disparity d found below it. It is the rectified counterpart of the reference's layer pyramid
 * (DenseStereoPyramid.Run, DenseStereoSgm.cs:328-353, where RangeFromPrevious :91-140 seeds each layer's
 * per-pixel range from the previous layer). base describes the full-resolution pair; base.disparities is the
 * plane count of the finest level (level k has base.disparities >> (levels-1-k)). 2*radius must be a
 * multiple of base.align and at most 128. Outputs are the finest level's: disp = start + first-minimum
 * index, sub = disp + parabola offset (each may be NULL). No left/right check. */
typedef struct {
    lux_sgm_rect_params base;
    int32_t levels;      /* 1..4 */
    int32_t radius[3];   /* levels 1 .. levels-1, coarse to fine */
} lux_sgm_pyramid_params;
int lux_sgm_pyramid_bytes(const lux_sgm_pyramid_params* p, uint64_t* bytes);
int lux_sgm_pyramid(lux_sgm_ctx* ctx, const lux_sgm_pyramid_params* p, const void* left, const void* right,
                    uint16_t* disp, float* sub);
int lux_sgm_pyramid_device(lux_sgm_ctx* ctx, const lux_sgm_pyramid_params* p, const void* d_left, const void* d_right,
                           uint16_t* d_disp, float* d_sub, void* stream);

/* Context options (no counterpart in the reference: execution strategy only, results are identical).
 * LUX_SGM_OPT_FUSED_BANDS: 1 runs the three vertical/diagonal paths of each pass fused in band tiles that
 * synchronise through device flags (up to 256 planes); 0 (default) runs one sweep kernel per path. */
#define LUX_SGM_OPT_FUSED_BANDS 1
/* LUX_SGM_OPT_LITERAL: 1 runs every layer of at most 1024 lanes (lux_sgm_compute_layer*, lux_sgm_dense_layer,
 * the dense pyramid) on the line-buffer-literal path: one wavefront per row follows the reference's tile
 * schedule against its own line buffers. Slower; layers that need it take it whatever this option says. */
#define LUX_SGM_OPT_LITERAL 2
int lux_sgm_set_option(lux_sgm_ctx* ctx, int32_t option, int32_t value);

/* Stage timing, the per-layer timing log of the reference's pyramid driver (DenseStereoSgm.cs:347-349)
 * at kernel granularity. With profiling on, each rectified request records HIP events on its stream
 * between launches; lux_sgm_stage_times waits for the last profiled request and writes the durations in
 * ms of its stages, in order: census codes (or the AD cost volume), pass +1 path -> (census cost fused),
 * pass +1 paths down-right, down, down-left (three stages; with LUX_SGM_OPT_FUSED_BANDS the first holds
 * the one fused launch and the other two are empty), the pass +1 border replay, pass -1 paths up-left, up,
 * up-right (likewise), pass -1 path <- (winner-take-all fused), the pass -1 border replay, and the
 * column-0 / validity finish. Returns the number of stages written (12), or < 0. */
int lux_sgm_set_profiling(lux_sgm_ctx* ctx, int on);
int lux_sgm_stage_times(lux_sgm_ctx* ctx, float* ms, int32_t cap);

/* Schedule plan inspection (host only, no device work): for the reference's wavefront over W x H with
 * tile T, the number of visits of the pass +1 border column (x = W-1) and of the pass -1 border column
 * (x = 0), and the seed row (the last pass +1 row written to line-buffer half 0, -1 if none). Returns
 * LUX_SGM_EUNSUPPORTED where the schedule falls outside what the device path reproduces (odd T). */
int lux_sgm_plan_info(int32_t width, int32_t height, int32_t tile, int32_t xs_clamp_w, int32_t* visits_plus,
                      int32_t* visits_minus, int32_t* seed_row);

/* The visit lists themselves: per visit 4 int32 (y, half, label of path 0, label of path 1); a label is
 * -1 for the initial 2000 fill, y' >= 0 for the neighbour column's state at row y' of the same pass, or
 * -2 - c for pass +1's leftover half-0 state of cell c. `cap` entries of room in each list. */
int lux_sgm_plan_visits(int32_t width, int32_t height, int32_t tile, int32_t xs_clamp_w, int32_t* plus, int32_t cap_plus,
                        int32_t* minus, int32_t cap_minus);

/* ---------------------------------------------------------------------------------------------------------
 * Batched form: all pairs of one capture (ExportBuild.cs:87-112 runs the dense stage once per capture) in one
 * graph launch. Every pointer of a pair is device memory; disp/sub/valid may each be NULL.
 */
typedef struct {
    lux_sgm_rect_params params;
    const void* left;
    const void* right;
    uint16_t* disp;
    float* sub;
    uint8_t* valid;
} lux_sgm_pair;

/* A batch schedule. executor = 0: each pair's launch sequence on one of `branches` concurrent graph branches,
 * longest first onto the least loaded branch; `bands` selects the fused-band paths (LUX_SGM_OPT_FUSED_BANDS).
 * executor = 1: the persistent executor, every stage of every pair in one launch (resident wavefronts pull work
 * items by ticket; an item waits for the stages it reads); executor = 7: the executor in one launch per pass group
 * (census codes, pass +1 row paths, pass +1 sweeps and border chains, pass +1 border replays, pass -1 sweeps and
 * chains, pass -1 row paths, pass -1 replays and finish). The executor takes census pairs whose path costs fit a
 * byte (up to 256 planes). order (executor): 0 stage by stage in batch order, 1 the pairs with the longest chains
 * first within each stage. ms: measured time of one batch under this schedule (0 when not measured). */
typedef struct {
    int32_t bands;
    int32_t branches;
    float ms;
    int32_t executor;
    int32_t order;
} lux_sgm_schedule;

/* Enqueue the n pairs on `stream` (NULL = the context's stream) as one hipGraph launch. The graph is built on
 * the first call with a given batch (descriptors and pointers) and replayed afterwards. The schedule is the
 * one lux_sgm_batch_tune chose for pairs of these shapes, else one branch per pair up to four. Returns once
 * enqueued; results are those of lux_sgm_rectified_device on each pair alone, bit for bit. No counterpart in
 * the reference (its dense stage runs one layer at a time, DenseStereoSgm.cs:346-351). */
int lux_sgm_batch(lux_sgm_ctx* ctx, int32_t n, const lux_sgm_pair* pairs, void* stream);

/* Use schedule `sch` for batches of these pairs' shapes (as if lux_sgm_batch_tune had chosen it). An executor
 * schedule for pairs the executor does not take fails with LUX_SGM_EUNSUPPORTED. */
int lux_sgm_batch_set_schedule(lux_sgm_ctx* ctx, int32_t n, const lux_sgm_pair* pairs, const lux_sgm_schedule* sch);

/* Autotuner: time each candidate schedule on this batch (each run once to warm, then `reps` timed runs) and
 * keep the fastest for batches of the same shapes. best receives it; tried (if not NULL) receives up to cap
 * candidates with their times. Returns the number of candidates tried, or < 0. */
int lux_sgm_batch_tune(lux_sgm_ctx* ctx, int32_t n, const lux_sgm_pair* pairs, void* stream, int32_t reps,
                       lux_sgm_schedule* best, lux_sgm_schedule* tried, int32_t cap);

/* ---------------------------------------------------------------------------------------------------------
 * Served form: a request loop over `slots` concurrent device slots (each a stream and a pooled workspace)
 * that admits depth requests in arrival order as slots and HBM budget free up, so that requests of
 * mixed sizes overlap on the device. It replaces the per-capture memory admission of the reference's job
 * queue (JobQueue.cs:23-31) and the 1 GiB layer rule (DenseStereoSgm.cs:330,339-344) for this stage.
 */
typedef struct lux_sgm_server lux_sgm_server;

#define LUX_SGM_MEM_HOST 0   /* images and outputs are host memory: inputs are copied at submit */
#define LUX_SGM_MEM_DEVICE 1 /* device memory, valid until the request completes */

#define LUX_SGM_REQ_PENDING 0
#define LUX_SGM_REQ_RUNNING 1
#define LUX_SGM_REQ_DONE 2

typedef struct {
    double queued_ms;   /* submit -> admitted to a slot */
    double run_ms;      /* admitted -> device work complete */
    double total_ms;    /* submit -> complete: the request's latency */
    int32_t slot;
    int32_t status;     /* LUX_SGM_OK or the error class */
} lux_sgm_request_stats;

typedef struct {
    uint64_t submitted, completed, failed;
    int32_t pending, running, slots, reserved;
    uint64_t budget, bytes_held, bytes_peak;
} lux_sgm_server_info;

/* hbm_budget bounds the device bytes of all slot workspaces together (0 = no bound); slots in 1..16. */
int lux_sgm_server_create(int device, uint64_t hbm_budget, int32_t slots, lux_sgm_server** out);

#define LUX_SGM_SERVER_DEVICE_LOOP 1u /* the request loop runs on the device (below) */

/* lux_sgm_server_create with flags. LUX_SGM_SERVER_DEVICE_LOOP: the loop that admits and schedules requests is a
 * persistent kernel rather than a host thread. Requests go through a ring in pinned host memory; the kernel admits
 * them in arrival order while one of `slots` is free and the request's buffers fit in a device pool of hbm_budget
 * bytes (required), runs the stages of every admitted request as work items, smallest request first, so their
 * work overlaps on the device, and marks each done in the ring. Host images and outputs are staged by the kernel
 * itself. It takes rectified census requests whose path costs fit a byte (up to 256 planes); others fail at
 * submit with LUX_SGM_EUNSUPPORTED, as do coarse-to-fine requests and cancel. Results equal lux_sgm_rectified's.
 * The kernel leaves after 20 ms with nothing to do and is launched again by the next submit. No counterpart in the
 * reference (its job queue runs on the host, Lux.Library/Core/JobQueue.cs:23-31). */
int lux_sgm_server_create_ex(int device, uint64_t hbm_budget, int32_t slots, uint32_t flags, lux_sgm_server** out);
/* Waits for running requests, cancels pending ones, frees everything. */
int lux_sgm_server_destroy(lux_sgm_server* srv);
const char* lux_sgm_server_last_error(const lux_sgm_server* srv);
/* Queue one rectified request. mem = LUX_SGM_MEM_HOST: left/right are copied before the call returns and
 * disp/sub/valid (each may be NULL) are written when lux_sgm_server_wait reports completion.
 * mem = LUX_SGM_MEM_DEVICE: every pointer is device memory that stays valid until completion. A request
 * larger than the whole budget fails here with LUX_SGM_ENOMEM. */
int lux_sgm_server_submit(lux_sgm_server* srv, const lux_sgm_rect_params* p, int32_t mem, const void* left,
                          const void* right, uint16_t* disp, float* sub, uint8_t* valid, uint64_t* ticket);
/* Queue one coarse-to-fine request (lux_sgm_pyramid_params); the same memory rules as lux_sgm_server_submit,
 * outputs disp and sub of the finest level. */
int lux_sgm_server_submit_pyramid(lux_sgm_server* srv, const lux_sgm_pyramid_params* p, int32_t mem, const void* left,
                                  const void* right, uint16_t* disp, float* sub, uint64_t* ticket);
/* State of a request (LUX_SGM_REQ_*), without waiting. */
int lux_sgm_server_poll(lux_sgm_server* srv, uint64_t ticket, int32_t* state);
/* Wait up to timeout_ms (< 0: no limit) for a request; on completion copy host outputs, fill stats (may be
 * NULL), forget the ticket and return the request's status. LUX_SGM_ECANCELED on timeout (the request
 * stays queued) or for a cancelled request. */
int lux_sgm_server_wait(lux_sgm_server* srv, uint64_t ticket, int32_t timeout_ms, lux_sgm_request_stats* stats);
/* Cancel a request that has not been admitted yet (the reference cancels between stages,
 * StereoAsyncApi.cs:51-63); a running request completes. */
int lux_sgm_server_cancel(lux_sgm_server* srv, uint64_t ticket);
int lux_sgm_server_info_get(lux_sgm_server* srv, lux_sgm_server_info* info);

#ifdef __cplusplus
}
#endif
#endif
