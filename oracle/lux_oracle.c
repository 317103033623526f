/*
 * lux_oracle.c -- CPU restatement of Lux.Engine's dense-stereo SGM layer.
 *
 * TEST INFRASTRUCTURE ONLY: this is the parity checker and the timed CPU baseline. The HIP product
 * library never links or calls it. Built with -ffp-contract=off so float expressions are evaluated
 * operation by operation, as the reference's JIT does (no fused multiply-add).
 *
 * Each function cites the reference lines it restates (paths relative to the mprovenz/lux checkout).
 */
#include "lux_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int mini(int a, int b) { return a < b ? a : b; }
static inline int maxi(int a, int b) { return a > b ? a : b; }

static inline uint16_t sat_add16(uint16_t a, uint16_t b) { uint32_t t = (uint32_t)a + b; return (uint16_t)(t > 65535u ? 65535u : t); }
static inline uint16_t sat_sub16(uint16_t a, uint16_t b) { return (uint16_t)(a > b ? a - b : 0); }
static inline uint16_t min16(uint16_t a, uint16_t b) { return a < b ? a : b; }
typedef uint16_t v8u __attribute__((vector_size(16)));
static inline v8u splat8(uint16_t v) { v8u r = {v, v, v, v, v, v, v, v}; return r; }
static inline v8u ld8(const uint16_t* p) { v8u r; memcpy(&r, p, 16); return r; }
static inline void st8(uint16_t* p, v8u v) { memcpy(p, &v, 16); }
static inline v8u vmin8(v8u a, v8u b) { v8u m = (v8u)(a < b); return (a & m) | (b & ~m); }
static inline v8u vadds8(v8u a, v8u b) { v8u t = a + b; return t | (v8u)(t < a); }
static inline v8u vsubs8(v8u a, v8u b) { v8u t = a - b; return t & (v8u)(a >= b); }
static int g_force_scalar = 0;
void lo_set_force_scalar(int v) { g_force_scalar = v; }

static inline int pix_at(const void* img, int img16, int stride, int x, int y)
{
    if (img16) return ((const uint16_t*)img)[(size_t)y * stride + x];
    return ((const uint8_t*)img)[(size_t)y * stride + x];
}

/* ---------------------------------------------------------------- census (project definition) */
int lo_census(const void* img, int img16, int W, int H, int stride, int radius, uint64_t* codes)
{
    if (!img || !codes || W <= 0 || H <= 0 || stride < W || radius < 1 || radius > 3) return -1;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int c = pix_at(img, img16, stride, x, y);
            uint64_t code = 0;
            int bit = 0;
            for (int dy = -radius; dy <= radius; dy++)
                for (int dx = -radius; dx <= radius; dx++) {
                    if (dx == 0 && dy == 0) continue;
                    int v = pix_at(img, img16, stride, clampi(x + dx, 0, W - 1), clampi(y + dy, 0, H - 1));
                    if (v < c) code |= (uint64_t)1 << bit;
                    bit++;
                }
            codes[(size_t)y * W + x] = code;
        }
    return 0;
}

/* ---------------------------------------------------------------- census cost volume */
int lo_cost_census(const uint64_t* cl, const uint64_t* cr, int W, int H, int D, int cap, uint8_t* raw, int threads)
{
    if (!cl || !cr || !raw || W <= 0 || H <= 0 || D <= 0 || cap < D || cap > 4096) return -1;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
#endif
    for (int y = 0; y < H; y++) /* row-parallel, DenseStereoSgm.cs:186-205 */
        for (int x = 0; x < W; x++) {
            uint8_t* o = raw + ((size_t)y * W + x) * cap;
            uint64_t a = cl[(size_t)y * W + x];
            const uint64_t* rrow = cr + (size_t)y * W;
            for (int d = 0; d < D; d++) o[d] = (uint8_t)__builtin_popcountll(a ^ rrow[maxi(x - d, 0)]);
            for (int d = D; d < cap; d++) o[d] = 255; /* DenseStereo.cs:380 */
        }
    return 0;
}

/* ---------------------------------------------------------------- truncated AD cost */
int lo_cost_ad(const void* L, const void* R, int img16, int bits, int W, int H, int stride, int D, int cap,
               int adcap, uint8_t* raw, int threads)
{
    if (!L || !R || !raw || W <= 0 || H <= 0 || stride < W || D <= 0 || cap < D || adcap < 0 || adcap > 28) return -1;
    int shift = img16 ? bits - 8 : 0;
    if (shift < 0 || shift > 8) return -1;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
#endif
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            uint8_t* o = raw + ((size_t)y * W + x) * cap;
            for (int d = 0; d < D; d++) {
                int xr = maxi(x - d, 0), s = 0;
                for (int dy = -1; dy <= 1; dy++) {
                    int yy = clampi(y + dy, 0, H - 1);
                    for (int dx = -1; dx <= 1; dx++) {
                        int a = pix_at(L, img16, stride, clampi(x + dx, 0, W - 1), yy);
                        int b = pix_at(R, img16, stride, clampi(xr + dx, 0, W - 1), yy);
                        int t = abs(a - b) >> shift;
                        s += t < adcap ? t : adcap;
                    }
                }
                o[d] = (uint8_t)s;
            }
            for (int d = D; d < cap; d++) o[d] = 255;
        }
    return 0;
}

/* ---------------------------------------------------------------- BoxFilter3, DenseStereo.cs:138-159 */
int lo_box3(const void* src, int img16, int W, int H, int maxval, void* dst)
{
    if (!src || !dst || W <= 0 || H <= 0) return -1;
    for (int y = 0; y < H; y++) {
        int y0 = maxi(y - 1, 0), y1 = mini(y + 1, H - 1), rows = y1 - y0 + 1;
        for (int x = 0; x < W; x++) {
            int x0 = maxi(x - 1, 0), x1 = mini(x + 1, W - 1), cols = x1 - x0 + 1;
            float inv = 1.0f / ((float)cols * (float)rows);
            int sum = 0;
            for (int yy = y0; yy <= y1; yy++)
                for (int xx = x0; xx <= x1; xx++) sum += pix_at(src, img16, W, xx, yy);
            float v = (float)sum * inv;
            int r = (int)rintf(v); /* cvtps2dq, round to nearest even */
            r = clampi(r, 0, maxval);
            if (img16) ((uint16_t*)dst)[(size_t)y * W + x] = (uint16_t)r;
            else ((uint8_t*)dst)[(size_t)y * W + x] = (uint8_t)r;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------- DenseLayer, DenseStereoSgm.cs:13-325 */
typedef struct {
    const lo_layer_params* p;
    int W, H, MinLo, MaxHi;
    uint16_t* cap;
    uint64_t* off;
    const uint8_t* raw;
    uint16_t* agg;
    /* line buffers, persistent across both passes (:28-29, AllocLineBuffers :158-165) */
    uint16_t *L, *Min, *RangeS, *RangeC;
    float* Pix; /* 2*slots*4 */
    long Llen;
    int slotStride, halfL, halfM, halfR;
    uint8_t* rawtmp;
    int fault;
} Layer;

/* the cubic of the fast exp2, highest power first (:168-169) */
static float f_of_bits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t bits_of_f(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

int lo_p2_adaptive(float p2Base, const float g[4], const float q[4], const float gw[4])
{
    const float E3 = f_of_bits(0x3d9fcb52), E2 = f_of_bits(0x3e677e26), E1 = f_of_bits(0x3f322226), E0 = f_of_bits(0x3f7ffb19);
    float d0 = (g[0] - q[0]) * gw[0], d1 = (g[1] - q[1]) * gw[1], d2 = (g[2] - q[2]) * gw[2];
    float S = (fabsf(d0) + fabsf(d2)) + fabsf(d1);
    float e = fmaxf(fminf(-S, 128.0f), -126.0f);
    int ei = (int)e + (e < 0.0f ? -1 : 0);
    float f = e - (float)ei;
    float poly = ((E3 * f + E2) * f + E1) * f + E0;
    float E = f_of_bits(bits_of_f(poly) + ((uint32_t)ei << 23));
    return (int)(p2Base * E);
}

static inline long LIdx(const Layer* s, int half, int slot, int path)
{
    return (long)half * s->halfL + (long)slot * s->slotStride + 1 + (long)path * (s->MaxHi + 1); /* :166 */
}
static inline int Rd(const Layer* s, long i) { return i < 0 ? 2000 : (i < s->Llen ? s->L[i] : 0); } /* :324 */

static void Pixel(Layer* s, int x, int y, int dir, int cur, int other)
{
    const lo_layer_params* P = s->p;
    int W = s->W;
    int pi = y * W + x, slot = x + 1;
    float g[4] = {0, 0, 0, 0};
    if (P->guidance)
        for (int c = 0; c < 4; c++) g[c] = (float)P->guidance[(size_t)pi * 4 + c]; /* :237-238 */
    memcpy(&s->Pix[((size_t)cur * s->halfR + slot) * 4], g, sizeof g);        /* :239 */
    int start = P->start[pi], count = P->count[pi], cap = s->cap[pi];
    const uint8_t* raw;
    if (P->skip && P->skip[pi] != 0) { /* skipped pixels: 0 in range, 255 padding (:242) */
        memset(s->rawtmp, 0, (size_t)cap);
        for (int k = count; k < cap; k++) s->rawtmp[k] = 255;
        raw = s->rawtmp;
    } else
        raw = s->raw + s->off[pi];
    int sp = s->RangeS[cur * s->halfR + slot], cp = s->RangeC[cur * s->halfR + slot];
    uint16_t* agg = s->agg + s->off[pi];
    int P1 = P->p1;
    float p2Base = (float)P->p1 * P->p2_scale;
    for (int p = 0; p < 4; p++) {
        long baseCur = LIdx(s, cur, slot, p);
        { /* stale cleanup of the slot's previously stored range (:261-264) */
            long q0 = sp - s->MinLo, k0 = q0 < 0 ? -q0 : 0, k1 = cp;
            if (s->Llen - baseCur - q0 < k1) k1 = s->Llen - baseCur - q0;
            for (long k = k0; k < k1; k++) s->L[baseCur + q0 + k] = 2000;
        }
        int ns = slot + (p == 2 ? 0 : p == 3 ? dir : -dir);
        int half = p == 0 ? cur : other; /* :265 */
        int minPrev = s->Min[half * s->halfM + ns * 4 + p];
        int P2;
        if (P->p2_const >= 0)
            P2 = P->p2_const;
        else { /* adaptive P2 (:267-273) */
            const float* q = &s->Pix[((size_t)half * s->halfR + ns) * 4];
            P2 = lo_p2_adaptive(p2Base, g, q, P->gw);
        }
        int P2p = mini(P2 + minPrev, 65535); /* :274 */
        if (cap == 0) { s->Min[cur * s->halfM + slot * 4 + p] = 0xffff; continue; }
        long nb = LIdx(s, half, ns, p);
        int runMin = 0xffff;
        long q = start - s->MinLo;
        /* vector form (:280-299): whole 8-lane blocks, every read and write inside the buffer, the written
         * range away from the range read, P2p >= 0. The loop below has no cross-lane dependence, so the
         * compiler vectorises it; it gives the scalar form's values under these conditions. */
        int vec = !g_force_scalar && (cap & 7) == 0 && P2p >= 0 && nb + q - 1 >= 0 && nb + q + cap + 1 <= s->Llen &&
                  baseCur + q >= 0 && baseCur + q + cap <= s->Llen && labs(baseCur - nb) >= cap + 16;
        if (vec) { /* eight u16 lanes per step, as Vector128<ushort> */
            const uint16_t* pn = s->L + nb + q;
            uint16_t* pw = s->L + baseCur + q;
            v8u vP1 = splat8((uint16_t)P1), vP2 = splat8((uint16_t)P2p), vMin = splat8((uint16_t)minPrev);
            v8u run = splat8(0xffff);
            for (int b = 0; b < cap; b += 8) {
                v8u lm = ld8(pn + b - 1), l0 = ld8(pn + b), lp = ld8(pn + b + 1);
                v8u m = vmin8(vmin8(vadds8(lm, vP1), l0), vmin8(vadds8(lp, vP1), vP2));
                v8u r16;
                for (int l = 0; l < 8; l++) r16[l] = raw[b + l];
                v8u lc = vsubs8(vadds8(r16, m), vMin);
                st8(pw + b, lc);
                run = vmin8(run, lc);
                st8(agg + b, vadds8(ld8(agg + b), lc));
            }
            uint16_t rm = 0xffff;
            for (int l = 0; l < 8; l++) rm = min16(rm, run[l]);
            s->Min[cur * s->halfM + slot * 4 + p] = rm;
            continue;
        }
        /* scalar form (:302-312) */
        for (int i = 0; i < cap; i++) {
            long qi = q + i;
            int Lm = Rd(s, nb + qi - 1), Lp = Rd(s, nb + qi + 1), L0 = Rd(s, nb + qi);
            int m = mini(mini(mini(Lm + P1, 65535), L0), mini(mini(Lp + P1, 65535), P2p));
            int lc = mini(raw[i] + m, 65535) - minPrev;
            if (lc < 0) lc = 0;
            long wi = baseCur + qi;
            if (wi < 0 || wi >= s->Llen) { s->fault = 1; return; } /* the reference throws here */
            s->L[wi] = (uint16_t)lc;
            if (lc < runMin) runMin = lc;
            agg[i] = (uint16_t)mini(agg[i] + lc, 65535);
        }
        s->Min[cur * s->halfM + slot * 4 + p] = (uint16_t)runMin;
    }
    s->RangeS[cur * s->halfR + slot] = (uint16_t)start;
    s->RangeC[cur * s->halfR + slot] = (uint16_t)cap;
}

int lo_layer_compute(const lo_layer_params* p, const uint8_t* raw_packed, uint16_t* agg_packed,
                     uint16_t* agg_pass1_packed, uint16_t* wta)
{
    if (!p || !raw_packed || !agg_packed || !p->start || !p->count || p->W <= 0 || p->H <= 0 || p->T <= 0 ||
        p->align <= 0 || p->p1 < 0 || p->p1 > 65535)
        return -1;
    if (p->p2_const < 0 && !p->guidance) return -1;
    Layer s;
    memset(&s, 0, sizeof s);
    s.p = p;
    s.W = p->W;
    s.H = p->H;
    int W = s.W, H = s.H, npx = W * H, align = p->align;
    s.cap = (uint16_t*)malloc(sizeof(uint16_t) * npx);
    s.off = (uint64_t*)malloc(sizeof(uint64_t) * (npx + 1));
    /* ranges: Init (:43-58); MinLo = min start, MaxHi = roundup(max(start+count), align) (:46, :51, :136) */
    int minLo = 65535, maxHi = 0, maxCap = 0;
    uint64_t o = 0;
    for (int i = 0; i < npx; i++) {
        int cap = ((p->count[i] + align - 1) / align) * align;
        s.cap[i] = (uint16_t)cap;
        s.off[i] = o;
        o += (uint64_t)cap;
        minLo = mini(minLo, p->start[i]);
        maxHi = maxi(maxHi, p->start[i] + p->count[i]);
        maxCap = maxi(maxCap, cap);
    }
    s.off[npx] = o;
    s.MinLo = minLo;
    s.MaxHi = ((maxHi + align - 1) / align) * align;
    s.raw = raw_packed;
    s.agg = agg_packed;
    memset(agg_packed, 0, sizeof(uint16_t) * o);
    /* AllocLineBuffers (:158-165) */
    int slots = W + 2;
    s.slotStride = 4 * s.MaxHi + 5;
    s.halfL = slots * s.slotStride;
    s.Llen = 2L * s.halfL;
    s.L = (uint16_t*)malloc(sizeof(uint16_t) * s.Llen);
    for (long i = 0; i < s.Llen; i++) s.L[i] = 2000;
    s.halfM = slots * 4;
    s.Min = (uint16_t*)malloc(sizeof(uint16_t) * 2 * s.halfM);
    for (int i = 0; i < 2 * s.halfM; i++) s.Min[i] = 2000;
    s.halfR = slots;
    s.RangeS = (uint16_t*)calloc(2 * slots, sizeof(uint16_t));
    s.RangeC = (uint16_t*)malloc(sizeof(uint16_t) * 2 * slots);
    for (int i = 0; i < 2 * slots; i++) s.RangeC[i] = (uint16_t)align;
    s.Pix = (float*)calloc((size_t)2 * slots * 4, sizeof(float));
    s.rawtmp = (uint8_t*)malloc((size_t)maxCap + 8);
    int T = p->T;
    /* Compute (:206-224): pass +1 then pass -1 over the sheared T x T wavefront tiles */
    for (int pass = 0; pass < 2 && !s.fault; pass++) {
        int dir = pass == 0 ? 1 : -1;
        int gwN = (W + T + T - 1) / T, ghN = (H + T - 1) / T;
        for (int j = 0; j < ghN && !s.fault; j++)
            for (int i = 0; i < gwN && !s.fault; i++) {
                int x0 = dir > 0 ? i * T : W - 1 - i * T, y0 = dir > 0 ? j * T : H - 1 - j * T;
                int yEnd = clampi(y0 + dir * T, -1, H);
                for (int r = 0;; r++) {
                    int y = y0 + dir * r;
                    if (y == yEnd) break;
                    int xs = clampi(x0 - dir * r, 0, (p->xsw && dir > 0) ? W : W - 1);
                    int xe = clampi(x0 + dir * (T - r), -1, W);
                    if (xs == xe) continue;
                    int cur = (r & 1) == 0 ? 1 : 0, other = 1 - cur; /* :219 */
                    for (int x = xs; x != xe && !s.fault; x += dir) Pixel(&s, x, y, dir, cur, other);
                }
            }
        if (dir > 0 && agg_pass1_packed) memcpy(agg_pass1_packed, agg_packed, sizeof(uint16_t) * o); /* :223 */
    }
    if (!s.fault && wta) /* first-minimum WTA (:225-231); the plane index start+arg */
        for (int i = 0; i < npx; i++) {
            const uint16_t* a = agg_packed + s.off[i];
            int cnt = p->count[i], best = 0xffff, arg = 0;
            for (int k = 0; k < cnt; k++)
                if (a[k] < best) { best = a[k]; arg = k; }
            wta[i] = (uint16_t)(p->start[i] + arg);
        }
    int fault = s.fault;
    free(s.cap); free(s.off); free(s.L); free(s.Min); free(s.RangeS); free(s.RangeC); free(s.Pix); free(s.rawtmp);
    return fault ? -2 : 0;
}

/* ---------------------------------------------------------------- WTA + subpixel + L/R */
int lo_wta_sub_lr(const uint16_t* agg, int W, int H, int D, int cap, int tol, uint16_t* disp, float* sub, uint8_t* valid)
{
    if (!agg || W <= 0 || H <= 0 || D <= 0 || cap < D) return -1;
    int* dr = (int*)malloc(sizeof(int) * W);
    int* dl = (int*)malloc(sizeof(int) * W);
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            const uint16_t* a = agg + ((size_t)y * W + x) * cap;
            int best = 0xffff, arg = 0;
            for (int d = 0; d < D; d++)
                if (a[d] < best) { best = a[d]; arg = d; }
            dl[x] = arg;
            if (disp) disp[(size_t)y * W + x] = (uint16_t)arg;
            if (sub) {
                float v = (float)arg;
                if (arg > 0 && arg < D - 1) {
                    int l = a[arg - 1], c = a[arg], r = a[arg + 1];
                    int den = l - 2 * c + r; /* exact in int and in float (< 2^24) */
                    if (den > 0) v = (float)arg + (0.5f * (float)(l - r)) / (float)den;
                }
                sub[(size_t)y * W + x] = v;
            }
        }
        for (int xr = 0; xr < W; xr++) {
            int best = 0xffff, arg = 0;
            for (int d = 0; d < D && xr + d < W; d++) {
                int v = agg[((size_t)y * W + xr + d) * cap + d];
                if (v < best) { best = v; arg = d; }
            }
            dr[xr] = arg;
        }
        if (valid)
            for (int x = 0; x < W; x++) {
                int xr = x - dl[x];
                valid[(size_t)y * W + x] = (uint8_t)(xr >= 0 && abs(dl[x] - dr[xr]) <= tol);
            }
    }
    free(dr);
    free(dl);
    return 0;
}

/* ---------------------------------------------------------------- whole rectified pipeline */
static double now_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int lo_rectified(const void* left, const void* right, int img16, int bits, int W, int H, int stride, int cost, int radius,
                 int adcap, int D, int align, int p1, int p2, int T, int xsw, int tol, int threads, uint16_t* disp,
                 float* sub, uint8_t* valid, uint16_t* agg, double* times_ms)
{
    if (!left || !right || W <= 0 || H <= 0 || D <= 0 || align <= 0 || p2 < 0) return -1;
    int cap = ((D + align - 1) / align) * align;
    size_t npx = (size_t)W * H, cells = npx * cap;
    uint8_t* raw = (uint8_t*)malloc(cells);
    uint16_t* ag = agg ? agg : (uint16_t*)malloc(cells * sizeof(uint16_t));
    uint16_t* start = (uint16_t*)calloc(npx, sizeof(uint16_t));
    uint16_t* count = (uint16_t*)malloc(npx * sizeof(uint16_t));
    if (!raw || !ag || !start || !count) return -3;
    for (size_t i = 0; i < npx; i++) count[i] = (uint16_t)D;
    int rc = 0;
    double t0 = now_ms();
    if (cost == 0) {
        uint64_t* cl = (uint64_t*)malloc(npx * 8);
        uint64_t* cr = (uint64_t*)malloc(npx * 8);
        rc = lo_census(left, img16, W, H, stride, radius, cl);
        if (!rc) rc = lo_census(right, img16, W, H, stride, radius, cr);
        if (!rc) rc = lo_cost_census(cl, cr, W, H, D, cap, raw, threads);
        free(cl);
        free(cr);
    } else
        rc = lo_cost_ad(left, right, img16, bits, W, H, stride, D, cap, adcap, raw, threads);
    double t1 = now_ms();
    if (!rc) {
        lo_layer_params lp;
        memset(&lp, 0, sizeof lp);
        lp.W = W; lp.H = H; lp.T = T; lp.align = align; lp.xsw = xsw; lp.p1 = p1; lp.p2_const = p2;
        lp.start = start; lp.count = count;
        rc = lo_layer_compute(&lp, raw, ag, NULL, NULL);
    }
    double t2 = now_ms();
    if (!rc) rc = lo_wta_sub_lr(ag, W, H, D, cap, tol, disp, sub, valid);
    double t3 = now_ms();
    if (times_ms) { times_ms[0] = t1 - t0; times_ms[1] = t2 - t1; times_ms[2] = t3 - t2; }
    free(raw);
    if (!agg) free(ag);
    free(start);
    free(count);
    return rc;
}
