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

#include <immintrin.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "rcp_table.h"
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

/* Init (:34-62) and AllocLineBuffers (:158-165) of one layer. */
static int layer_init(Layer* s, const lo_layer_params* p, const uint8_t* raw_packed, uint16_t* agg_packed)
{
    if (!p || !raw_packed || !agg_packed || !p->start || !p->count || p->W <= 0 || p->H <= 0 || p->T <= 0 ||
        p->align <= 0 || p->p1 < 0 || p->p1 > 65535)
        return -1;
    if (p->p2_const < 0 && !p->guidance) return -1;
    memset(s, 0, sizeof *s);
    s->p = p;
    s->W = p->W;
    s->H = p->H;
    int W = s->W, H = s->H, npx = W * H, align = p->align;
    s->cap = (uint16_t*)malloc(sizeof(uint16_t) * npx);
    s->off = (uint64_t*)malloc(sizeof(uint64_t) * (npx + 1));
    /* ranges: Init (:43-58); MinLo = min start, MaxHi = roundup(max(start+count), align) (:46, :51, :136) */
    int minLo = 65535, maxHi = 0, maxCap = 0;
    uint64_t o = 0;
    for (int i = 0; i < npx; i++) {
        int cap = ((p->count[i] + align - 1) / align) * align;
        s->cap[i] = (uint16_t)cap;
        s->off[i] = o;
        o += (uint64_t)cap;
        minLo = mini(minLo, p->start[i]);
        maxHi = maxi(maxHi, p->start[i] + p->count[i]);
        maxCap = maxi(maxCap, cap);
    }
    s->off[npx] = o;
    s->MinLo = minLo;
    s->MaxHi = ((maxHi + align - 1) / align) * align;
    s->raw = raw_packed;
    s->agg = agg_packed;
    memset(agg_packed, 0, sizeof(uint16_t) * o);
    /* AllocLineBuffers (:158-165) */
    int slots = W + 2;
    s->slotStride = 4 * s->MaxHi + 5;
    s->halfL = slots * s->slotStride;
    s->Llen = 2L * s->halfL;
    s->L = (uint16_t*)malloc(sizeof(uint16_t) * s->Llen);
    for (long i = 0; i < s->Llen; i++) s->L[i] = 2000;
    s->halfM = slots * 4;
    s->Min = (uint16_t*)malloc(sizeof(uint16_t) * 2 * s->halfM);
    for (int i = 0; i < 2 * s->halfM; i++) s->Min[i] = 2000;
    s->halfR = slots;
    s->RangeS = (uint16_t*)calloc(2 * slots, sizeof(uint16_t));
    s->RangeC = (uint16_t*)malloc(sizeof(uint16_t) * 2 * slots);
    for (int i = 0; i < 2 * slots; i++) s->RangeC[i] = (uint16_t)align;
    s->Pix = (float*)calloc((size_t)2 * slots * 4, sizeof(float));
    s->rawtmp = (uint8_t*)malloc((size_t)maxCap + 8);
    (void)H;
    return 0;
}

/* first-minimum WTA (:225-231); the plane index start+arg */
static void layer_wta(const Layer* s, uint16_t* wta)
{
    const lo_layer_params* p = s->p;
    for (int i = 0; i < s->W * s->H; i++) {
        const uint16_t* a = s->agg + s->off[i];
        int cnt = p->count[i], best = 0xffff, arg = 0;
        for (int k = 0; k < cnt; k++)
            if (a[k] < best) { best = a[k]; arg = k; }
        wta[i] = (uint16_t)(p->start[i] + arg);
    }
}

static int layer_finish(Layer* s, uint16_t* wta)
{
    if (!s->fault && wta) layer_wta(s, wta);
    int fault = s->fault;
    free(s->cap); free(s->off); free(s->L); free(s->Min); free(s->RangeS); free(s->RangeC); free(s->Pix); free(s->rawtmp);
    return fault ? -2 : 0;
}

int lo_layer_compute(const lo_layer_params* p, const uint8_t* raw_packed, uint16_t* agg_packed,
                     uint16_t* agg_pass1_packed, uint16_t* wta)
{
    Layer s;
    if (layer_init(&s, p, raw_packed, agg_packed)) return -1;
    int W = s.W, H = s.H, T = p->T;
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
        if (dir > 0 && agg_pass1_packed) memcpy(agg_pass1_packed, agg_packed, sizeof(uint16_t) * s.off[W * H]); /* :223 */
    }
    return layer_finish(&s, wta);
}

int lo_layer_compute_visits(const lo_layer_params* p, const uint8_t* raw_packed, uint16_t* agg_packed,
                            uint16_t* agg_pass1_packed, uint16_t* wta, const int32_t* visits, int32_t n_plus,
                            int32_t n_minus)
{
    Layer s;
    if (!visits || n_plus < 0 || n_minus < 0 || layer_init(&s, p, raw_packed, agg_packed)) return -1;
    for (int k = 0; k < n_plus + n_minus && !s.fault; k++) {
        const int32_t* v = visits + 3 * (size_t)k;
        if (v[0] < 0 || v[0] >= s.W || v[1] < 0 || v[1] >= s.H || (v[2] & ~1)) { s.fault = 1; break; }
        Pixel(&s, v[0], v[1], k < n_plus ? 1 : -1, v[2], 1 - v[2]);
        if (k + 1 == n_plus && agg_pass1_packed) memcpy(agg_pass1_packed, agg_packed, sizeof(uint16_t) * s.off[s.W * s.H]);
    }
    return layer_finish(&s, wta);
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


/* ---------------------------------------------------------------- DenseUpsampleLayer */
uint32_t lo_rcp_bits(uint32_t b) /* IntelApprox.cs:21-29 */
{
    const uint32_t sign = b & 0x80000000u, e = (b >> 23) & 0xffu, m = b & 0x7fffffu;
    if (e == 0xffu) return m ? (b | 0x400000u) : sign;
    if (e == 0) return sign | 0x7f800000u;
    const uint32_t r = kRcpTable[m >> 11];
    const int re = (int)((r >> 23) & 0xffu) - ((int)e - 127);
    if (re <= 0) return sign;
    return sign | ((uint32_t)re << 23) | (r & 0x7fffffu);
}

uint32_t lo_rcpss_hw(uint32_t b)
{
    float x = f_of_bits(b);
    return bits_of_f(_mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(x))));
}

static float lo_rcp(float x) { return f_of_bits(lo_rcp_bits(bits_of_f(x))); }

/* The cubic tap (:58-63): Catmull-Rom written with the constants 9, -15, 6 and -3, 15, -24, 12, times 1/6. */
static float cubic_tap(float x)
{
    const float sixth = f_of_bits(0x3e2aaaab);
    float x2 = x * x, x3 = x2 * x;
    if (x < 1.0f) { float a = x3 * 9.0f; float b = x2 * -15.0f; b = b + 6.0f; b = b + a; return b * sixth; }
    if (x < 2.0f) { float a = x3 * -3.0f; float b = x2 * 15.0f; float c = x * -24.0f; c = c + 12.0f; c = c + b; c = c + a; return c * sixth; }
    return 0.0f;
}

int lo_image_shift(const uint8_t* src, int W, int H, float sx, float sy, uint8_t* dst) /* :70-104 */
{
    if (!src || !dst || W <= 0 || H <= 0) return -1;
    const int ix = (int)floorf(sx), iy = (int)floorf(sy);
    const float tx = sx - (float)ix, ty = sy - (float)iy;
    const float kx[4] = {cubic_tap(tx + 1.0f), cubic_tap(tx), cubic_tap(1.0f - tx), cubic_tap(2.0f - tx)};
    const float ky[4] = {cubic_tap(ty + 1.0f), cubic_tap(ty), cubic_tap(1.0f - ty), cubic_tap(2.0f - ty)};
    float* row = (float*)malloc(sizeof(float) * (size_t)(W + 4) * 4);
    for (int y = 0; y < H; y++) {
        int rows[4];
        for (int i = 0; i < 4; i++) rows[i] = clampi(y + iy - 1 + i, 0, H - 1);
        for (int c = -2; c < W + 2; c++) { /* columns over all four source rows, clamp to edge */
            const int cx = clampi(c, 0, W - 1);
            for (int ch = 0; ch < 4; ch++) {
                float acc = (float)src[((size_t)rows[0] * W + cx) * 4 + ch] * ky[0];
                for (int i = 1; i < 4; i++) acc = (float)src[((size_t)rows[i] * W + cx) * 4 + ch] * ky[i] + acc;
                row[(c + 2) * 4 + ch] = acc;
            }
        }
        for (int x = 0; x < W; x++) {
            const int b = (x + ix - 1 + 2) * 4;
            for (int ch = 0; ch < 4; ch++) {
                const float v0 = row[b + ch], v1 = row[b + 4 + ch], v2 = row[b + 8 + ch], v3 = row[b + 12 + ch];
                const float s = (v3 * kx[3] + v2 * kx[2]) + (v1 * kx[1] + v0 * kx[0]);
                const int q = (int)rintf(s); /* round to nearest even */
                dst[((size_t)y * W + x) * 4 + ch] = (uint8_t)clampi(q, 0, 255);
            }
        }
    }
    free(row);
    return 0;
}

int lo_nearest_resize(const uint8_t* src, int sw, int sh, int W, int H, uint8_t* dst) /* :108-128 */
{
    if (!src || !dst || W <= 0 || H <= 0 || sw <= 0 || sh <= 0) return -1;
    const int stepY = (int)((double)sh * 65536.0 / (double)H), stepX = (int)((double)sw * 65536.0 / (double)W);
    int yacc = 0;
    for (int y = 0; y < H; y++) {
        const int syy = yacc >> 16;
        int xacc = 0;
        for (int x = 0; x < W; x++) {
            const int sxx = xacc >> 16;
            memcpy(dst + ((size_t)y * W + x) * 4, src + ((size_t)syy * sw + sxx) * 4, 4);
            xacc += stepX;
        }
        yacc += stepY;
    }
    return 0;
}

/* The bilateral weight's fast 2^t (:144-152): n = (int)t + (bits(t) >> 31), so -0 takes n = -1 and f = 1. */
static float exp2_fast_signbit(float t)
{
    const float c1 = f_of_bits(0x3d9fcb52), c2 = f_of_bits(0x3e677e26), c3 = f_of_bits(0x3f322226), c4 = f_of_bits(0x3f7ffb19);
    t = t < 128.0f ? t : 128.0f;
    t = t > -126.0f ? t : -126.0f;
    const int n = (int)t + ((int32_t)bits_of_f(t) >> 31);
    const float f = t - (float)n;
    const float p = ((f * c1 + c2) * f + c3) * f + c4;
    return f_of_bits(bits_of_f(p) + ((uint32_t)n << 23));
}

static float spatial_w(float f, int i) /* :213-218 */
{
    float d = fabsf(f - (float)(i * 2));
    d = d < 1.0f ? d : 1.0f;
    return (int)d == 0 ? 1.0f : f_of_bits(0x3eaaaaab);
}

int lo_bilateral_upsample(const float* lo, int lw, int lh, const uint8_t* guide, int W, int H, const uint8_t* g2, float sigma,
                          float* out) /* :156-210 */
{
    if (!lo || !guide || !g2 || !out || lw <= 0 || lh <= 0 || W <= 0 || H <= 0) return -1;
    const float k = 0.5f / (sigma * sigma), eps = f_of_bits(0x322bcc77);
    const int wm1 = lw - 1, hm1 = lh - 1;
    for (int y = 0; y < H; y++) {
        const int iy0 = (y - 1) / 2; /* C division, toward zero */
        const float fy = (float)y + -0.5f;
        for (int x = 0; x < W; x++) {
            const uint8_t* gp = guide + ((size_t)y * W + x) * 4;
            const float hr = gp[0], hg = gp[1], hb = gp[2];
            const int ix0 = (x - 1) / 2;
            const float fx = (float)x + -0.5f;
            const int ibx = clampi(ix0, 0, wm1), ix1 = clampi(ix0 + 1, 0, wm1);
            const float wx0 = spatial_w(fx, ibx), wx1 = spatial_w(fx, ix1);
            float accV = 0.0f, accS = 0.0f;
            for (int j = 0; j < 2; j++) {
                const int iy = clampi(iy0 + j, 0, hm1);
                const float wy = spatial_w(fy, iy);
                const uint8_t* q0 = g2 + ((size_t)iy * lw + ibx) * 4;
                float d0 = hr - q0[0], d1 = hg - q0[1], d2 = hb - q0[2];
                float sq = d1 * d1 + (d2 * d2 + d0 * d0);
                const float w0 = (exp2_fast_signbit(-(sq * k)) + eps) * (wy * wx0);
                const uint8_t* q1 = g2 + ((size_t)iy * lw + ix1) * 4;
                d0 = hr - q1[0]; d1 = hg - q1[1]; d2 = hb - q1[2];
                sq = d1 * d1 + (d2 * d2 + d0 * d0);
                const float w1 = (exp2_fast_signbit(-(sq * k)) + eps) * (wy * wx1);
                const float v0 = lo[(size_t)iy * lw + ibx] * w0, v1 = lo[(size_t)iy * lw + ix1] * w1;
                accV = v1 + (v0 + accV);
                accS = w1 + (w0 + accS);
            }
            out[(size_t)y * W + x] = accV / accS;
        }
    }
    return 0;
}

int lo_upsample_layer(const float* prev_depth, int pw, int ph, const uint8_t* guide, int W, int H, int setW, int setH,
                      float sigma, float* depth) /* :38-52 */
{
    if (!prev_depth || !guide || !depth || pw <= 0 || ph <= 0 || W <= 0 || H <= 0) return -1;
    if (W != setW * 2 || H != setH * 2) return -3;
    uint8_t* shifted = (uint8_t*)malloc((size_t)W * H * 4);
    uint8_t* g2 = (uint8_t*)malloc((size_t)pw * ph * 4);
    float* inv = (float*)malloc(sizeof(float) * (size_t)pw * ph);
    lo_image_shift(guide, W, H, 0.5f, 0.5f, shifted);
    lo_nearest_resize(shifted, W, H, pw, ph, g2);
    for (size_t i = 0; i < (size_t)pw * ph; i++) inv[i] = lo_rcp(prev_depth[i]);
    lo_bilateral_upsample(inv, pw, ph, guide, W, H, g2, sigma, depth);
    for (size_t i = 0; i < (size_t)W * H; i++) depth[i] = lo_rcp(depth[i]);
    free(shifted);
    free(g2);
    free(inv);
    return 0;
}


/* ======================================================================================================
 * Plane-sweep matching cost, DenseStereo.cs:217-381, and the layer set-up around it (DenseStereoSgm.cs).
 * ====================================================================================================== */

/* float -> int as the reference's runtime converts (saturating; NaN -> 0). */
static int f2i_sat(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f < -2147483648.0f) return INT32_MIN;
    return (int)f;
}

static inline int avg_u8(int a, int b) { return (a + b + 1) >> 1; } /* pavgb */

/* One camera's cost of one plane with the 4x4 target block at (ix-1, iy-1) (:276-291, :332-367): rows averaged with
 * the next row for a half-pixel y, then pixels with the next pixel for a half-pixel x; capped differences against the
 * reference block summed per channel over the 3x3 pixels; weighted and clamped to 65535. */
static int block_cost(const lo_view* v, const uint8_t refb[9][4], const uint8_t caps[4], int ix, int iy, int hx, int hy)
{
    int S[4] = {0, 0, 0, 0};
    for (int q = 0; q < 3; q++) {
        const uint8_t* r0 = v->rgba + ((size_t)(iy - 1 + q) * v->w + (ix - 1)) * 4;
        const uint8_t* r1 = r0 + (size_t)v->w * 4;
        for (int p = 0; p < 3; p++)
            for (int ch = 0; ch < 4; ch++) {
                int a = r0[p * 4 + ch], b = r0[(p + 1) * 4 + ch];
                if (hy) {
                    a = avg_u8(a, r1[p * 4 + ch]);
                    b = avg_u8(b, r1[(p + 1) * 4 + ch]);
                }
                const int t = hx ? avg_u8(a, b) : a;
                const int rb = refb[q * 3 + p][ch];
                int d = t > rb ? t - rb : rb - t;
                if (d > caps[ch]) d = caps[ch];
                S[ch] += d;
            }
    }
    const int cY = (S[0] * v->wq[0] + 16) >> 5, cU = (S[1] * v->wq[1] + 16) >> 5, cV = (S[2] * v->wq[2] + 16) >> 5,
              cA = (S[3] * v->wq[3] + 16) >> 5;
    const int ck = (cY + cU) + (cV + cA);
    return ck < 65535 ? ck : 65535;
}

int lo_matching_cost(const lo_sweep* s, int x, int y, int start, int count, uint16_t* cost)
{
    if (!s || !cost || start < 0 || count < 0 || start + count > s->n) return -1;
    for (int k = 0; k < s->nv; k++)
        if (s->views[k].w < 4 || s->views[k].h < 4) return -1;
    uint8_t refb[9][4];
    for (int r = -1; r <= 1; r++)
        for (int c = -1; c <= 1; c++) {
            const int yy = clampi(y + r, 0, s->rh - 1), xx = clampi(x + c, 0, s->rw - 1);
            memcpy(refb[(r + 1) * 3 + (c + 1)], s->ref + ((size_t)yy * s->rw + xx) * 4, 4);
        }
    memset(cost, 0, sizeof(uint16_t) * (size_t)count);
    const float xf = (float)x, yf = (float)y;
    for (int k = 0; k < s->nv; k++) {
        const lo_view* v = &s->views[k];
        const float* m = v->warp;
        const int x0 = 1, y0 = 1, x1 = v->w - 2, y1 = v->h - 2; /* rect + {1, 1, -2, -2} */
        int lix = 0, liy = 0, lhx = 0, lhy = 0, last = 0, have = 0;
        for (int d = 0; d < count; d++) {
            const float z = s->planes[start + d];
            const float xz = xf * z, yz = yf * z;
            const float px = ((z * m[8] + m[12]) + xz * m[0]) + yz * m[4];
            const float py = ((z * m[9] + m[13]) + xz * m[1]) + yz * m[5];
            const float pz = ((z * m[10] + m[14]) + xz * m[2]) + yz * m[6];
            const float inv = 1.0f / pz;
            const float fx = px * inv + 0.25f, fy = py * inv + 0.25f;
            int ix, iy, hx, hy;
            if (d == 0) { /* the first plane is clamped into the rect */
                const float cx = fminf(fmaxf(fx, (float)x0), (float)(x1 - 1)), cy = fminf(fmaxf(fy, (float)y0), (float)(y1 - 1));
                ix = f2i_sat(cx);
                iy = f2i_sat(cy);
                hx = f2i_sat(cx + cx) & 1;
                hy = f2i_sat(cy + cy) & 1;
            } else {
                ix = f2i_sat(fx);
                iy = f2i_sat(fy);
                hx = f2i_sat(fx + fx) & 1;
                hy = f2i_sat(fy + fy) & 1;
                /* the same position as the last costed plane, or outside the rect: the last cost again */
                if ((have && ix == lix && iy == liy && hx == lhx && hy == lhy) || ix < x0 || ix >= x1 || iy < y0 || iy >= y1) {
                    cost[d] = (uint16_t)(cost[d] + last);
                    continue;
                }
            }
            const int c = block_cost(v, (const uint8_t(*)[4])refb, s->caps, ix, iy, hx, hy);
            cost[d] = (uint16_t)(cost[d] + c);
            lix = ix; liy = iy; lhx = hx; lhy = hy; last = c; have = 1;
        }
    }
    return 0;
}

void lo_normalise(const uint16_t* cost, int count, int cap, int nv, uint8_t* raw)
{
    const float f = f_of_bits(0x3d17b426) / (float)nv;
    for (int d = 0; d < count; d++) raw[d] = (uint8_t)f2i_sat((float)cost[d] * f);
    for (int d = count; d < cap; d++) raw[d] = 255;
}

int lo_layer_sweep_cost(const lo_sweep* s, int W, int H, int scale, int align, const uint16_t* start, const uint16_t* count,
                        const uint8_t* skip, uint8_t* raw, int threads)
{
    if (!s || !start || !count || !raw || W <= 0 || H <= 0 || scale <= 0 || align <= 0 || s->nv <= 0) return -1;
    const size_t npx = (size_t)W * H;
    size_t* off = (size_t*)malloc(sizeof(size_t) * (npx + 1));
    off[0] = 0;
    for (size_t i = 0; i < npx; i++) off[i + 1] = off[i] + (size_t)((count[i] + align - 1) / align) * align;
    const int half = scale / 2;
    int bad = 0;
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) reduction(| : bad)
    for (int y = 0; y < H; y++) {
        uint16_t* cost = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(s->n + 1));
        for (int x = 0; x < W; x++) {
            const size_t i = (size_t)y * W + x;
            const int cap = (int)(off[i + 1] - off[i]), cnt = count[i];
            uint8_t* r = raw + off[i];
            if (skip && skip[i]) {
                memset(r, 0, (size_t)cnt);
                memset(r + cnt, 255, (size_t)(cap - cnt));
                continue;
            }
            const int xi = scale * x + half < s->rw - 1 ? scale * x + half : s->rw - 1;
            const int yi = scale * y + half < s->rh - 1 ? scale * y + half : s->rh - 1;
            if (lo_matching_cost(s, xi, yi, start[i], cnt, cost)) { bad = 1; continue; }
            lo_normalise(cost, cnt, cap, s->nv, r);
        }
        free(cost);
    }
    free(off);
    return bad ? -1 : 0;
}

int lo_guidance(const uint8_t* ref, int rw, int rh, int w, int h, int scale, uint8_t* out)
{
    if (!ref || !out || w <= 0 || h <= 0 || scale <= 0) return -1;
    if (rw == w && rh == h) {
        memcpy(out, ref, (size_t)w * h * 4);
        return 0;
    }
    const int half = scale / 2;
    for (int y = 0; y < h; y++) {
        const int sy = scale * y + half < rh - 1 ? scale * y + half : rh - 1;
        for (int x = 0; x < w; x++) {
            const int sx = scale * x + half < rw - 1 ? scale * x + half : rw - 1;
            memcpy(out + ((size_t)y * w + x) * 4, ref + ((size_t)sy * rw + sx) * 4, 4);
        }
    }
    return 0;
}

int lo_range_from_previous(const float* planes, int n, const float* prev_depth, const uint8_t* prev_skip, int pw, int ph,
                           int w, int h, int margin, uint16_t* start, uint16_t* count, int* min_lo, int* max_hi)
{
    if (!planes || !prev_depth || !start || !count || n <= 0 || pw <= 0 || ph <= 0 || w <= 0 || h <= 0) return -1;
    /* plane boundaries in inverse depth: midpoints of neighbouring planes' rcpss, the last one FLT_MAX */
    float* b = (float*)malloc(sizeof(float) * (size_t)n);
    for (int i = 0; i < n - 1; i++) b[i] = (lo_rcp(planes[i + 1]) + lo_rcp(planes[i])) * 0.5f;
    b[n - 1] = f_of_bits(0x7f7fffff);
    const size_t np = (size_t)pw * ph;
    uint16_t* idx = (uint16_t*)malloc(2 * np);
    for (size_t i = 0; i < np; i++) {
        const float v = lo_rcp(prev_depth[i]);
        int c = 0;
        while (c < n && b[c] <= v) c++; /* the leading run of boundaries at or below v */
        idx[i] = (uint16_t)c;
    }
    /* 4-wide min/max over x (-1..+2, clamped) of computed pixels, then over y (-1..+2, clamped) */
    uint16_t *hmin = (uint16_t*)malloc(2 * np), *hmax = (uint16_t*)malloc(2 * np), *vmin = (uint16_t*)malloc(2 * np),
             *vmax = (uint16_t*)malloc(2 * np);
    for (int y = 0; y < ph; y++)
        for (int x = 0; x < pw; x++) {
            int mn = 0xffff, mx = 0;
            for (int k = -1; k <= 2; k++) {
                const int xx = clampi(x + k, 0, pw - 1);
                if (prev_skip && prev_skip[(size_t)y * pw + xx]) continue;
                const int v = idx[(size_t)y * pw + xx];
                if (v < mn) mn = v;
                if (v > mx) mx = v;
            }
            hmin[(size_t)y * pw + x] = (uint16_t)mn;
            hmax[(size_t)y * pw + x] = (uint16_t)mx;
        }
    for (int y = 0; y < ph; y++)
        for (int x = 0; x < pw; x++) {
            int mn = 0xffff, mx = 0;
            for (int k = -1; k <= 2; k++) {
                const int yy = clampi(y + k, 0, ph - 1);
                const int a = hmin[(size_t)yy * pw + x], c = hmax[(size_t)yy * pw + x];
                if (a < mn) mn = a;
                if (c > mx) mx = c;
            }
            vmin[(size_t)y * pw + x] = (uint16_t)mn;
            vmax[(size_t)y * pw + x] = (uint16_t)mx;
        }
    const float fy = (float)(ph - 1) / (float)(h - 1), fx = (float)(pw - 1) / (float)(w - 1);
    int mhi = 0, mlo = INT32_MAX;
    for (int y = 0; y < h; y++) {
        const int py = f2i_sat((float)y * fy);
        for (int x = 0; x < w; x++) {
            const int px = f2i_sat((float)x * fx);
            int lo = vmin[(size_t)py * pw + px] - margin, hi = vmax[(size_t)py * pw + px] + 1;
            if (lo < 0) lo = 0;
            if (hi > n - 1) hi = n - 1;
            const size_t i = (size_t)y * w + x;
            start[i] = (uint16_t)lo;
            count[i] = (uint16_t)(hi - lo);
            if ((hi & 0xffff) > mhi) mhi = hi & 0xffff;
            if ((lo & 0xffff) < mlo) mlo = lo & 0xffff;
        }
    }
    if (min_lo) *min_lo = mlo;
    if (max_hi) *max_hi = mhi;
    free(b); free(idx); free(hmin); free(hmax); free(vmin); free(vmax);
    return 0;
}

int lo_layer_depth(const float* planes, int n, int W, int H, int align, const uint16_t* start, const uint16_t* count,
                   const uint16_t* agg, float* depth)
{
    if (!planes || !start || !count || !agg || !depth || align <= 0) return -1;
    size_t off = 0;
    for (size_t i = 0; i < (size_t)W * H; i++) {
        const int cnt = count[i], cap = ((cnt + align - 1) / align) * align;
        int best = 0xffff, arg = 0;
        for (int k = 0; k < cnt; k++)
            if (agg[off + k] < best) { best = agg[off + k]; arg = k; }
        const int pi = (uint16_t)(start[i] + arg);
        depth[i] = pi < n ? planes[pi] : 0.0f;
        off += (size_t)cap;
    }
    return 0;
}


/* ---- SkipMask and its generator ---- */
/* MT19937 (Matsumoto & Nishimura 1998): 624 words of state, the twist and the tempering of the published algorithm. */
typedef struct {
    uint32_t mt[624];
    int i;
} lo_mt;

static void mt_seed(lo_mt* g, uint32_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 624; i++) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->i = 624;
}

static uint32_t mt_next(lo_mt* g)
{
    if (g->i >= 624) {
        for (int k = 0; k < 624; k++) {
            const uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
            g->mt[k] = g->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->i = 0;
    }
    uint32_t y = g->mt[g->i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

uint32_t lo_mt19937(uint32_t seed, int n)
{
    lo_mt g;
    mt_seed(&g, seed);
    uint32_t v = 0;
    for (int k = 0; k < n; k++) v = mt_next(&g);
    return v;
}

int lo_skip_mask(int mode, int w, int h, uint8_t* out)
{
    if (w <= 0 || h <= 0 || !out) return -1;
    if (mode == 0) {
        memset(out, 0, (size_t)w * h);
        return 0;
    }
    if (mode != 2) return -2;
    memset(out, 0xff, (size_t)w * h);
    int nx = w / 64 + (64 < 2 * (w % 64) ? 1 : 0), ny = h / 64 + (64 < 2 * (h % 64) ? 1 : 0);
    if (nx < 1) nx = 1;
    if (ny < 1) ny = 1;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int x0 = 64 * i, y0 = 64 * j;
            const int x1 = w < x0 + 64 * (i == nx - 1 ? 2 : 1) ? w : x0 + 64 * (i == nx - 1 ? 2 : 1);
            const int y1 = h < y0 + 64 * (j == ny - 1 ? 2 : 1) ? h : y0 + 64 * (j == ny - 1 ? 2 : 1);
            lo_mt g;
            mt_seed(&g, 5489u);
            for (int y = y0; y < y1; y += 2)
                for (int x = x0; x < x1; x += 2) {
                    const int dy = (int)(mt_next(&g) & 1u), dx = (int)(mt_next(&g) & 1u);
                    const int yy = y + dy, xx = x + dx;
                    if (xx < w && yy < h) out[(size_t)yy * w + xx] = 0;
                }
        }
    return 0;
}

/* ---- layer set-up arithmetic of the L16 path (host, once per capture) ---- */
#include "rsqrt_table.h"

uint32_t lo_rsqrt_bits(uint32_t b) /* IntelApprox.cs:32-44 */
{
    const uint32_t sign = b & 0x80000000u, e = (b >> 23) & 0xffu, m = b & 0x7fffffu;
    if (e == 0xffu && m != 0) return b | 0x400000u; /* NaN -> quiet NaN */
    if (e == 0) return sign | 0x7f800000u;          /* +-0, denormals -> +-inf */
    if (sign != 0) return 0xffc00000u;               /* negative -> indefinite QNaN */
    if (e == 0xffu) return 0;                        /* +inf -> +0 */
    const int k = (int)e - 127;
    uint32_t r;
    int re;
    if ((k & 1) == 0) { r = kRsqrtOddTable[m >> 11]; re = (int)((r >> 23) & 0xffu) - k / 2; }
    else { r = kRsqrtEvenTable[m >> 11]; re = (int)((r >> 23) & 0xffu) - (k - 1) / 2; }
    if (re <= 0) return 0;
    return ((uint32_t)re << 23) | (r & 0x7fffffu);
}

uint32_t lo_rsqrtss_hw(uint32_t b)
{
    float x = f_of_bits(b);
    return bits_of_f(_mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(x))));
}

float lo_homography_sqrt(float x) /* Homography.cs:18 */
{
    float r = f_of_bits(lo_rsqrt_bits(bits_of_f(x)));
    float A = x * r, B = A * (-0.5f), C = A * r, D = C + (-3.0f), s = D * B;
    return x == 0.0f ? 0.0f : s;
}

static float lo_rcp_nr(float a) /* DenseStereo.cs:36-37 */
{
    float x0 = lo_rcp(a);
    return x0 + x0 * (1.0f - a * x0);
}

int lo_planes(float near_z, float far_z, const float* calibs, int ncal, float density, int align, float* planes, int cap,
              int* n_out) /* DenseStereo.cs:41-88; calibs: ncal records of K[9], T[3], R[9] */
{
    if (!(near_z < far_z)) return -1;
    if (!(near_z > 0.0f && far_z > 0.0f)) return -2;
    if (ncal <= 0 || align <= 0) return -4;
    float C[64][3];
    if (ncal > 64) return -4;
    for (int k = 0; k < ncal; k++) {
        const float* R = calibs + 21 * k + 12;
        const float* t = calibs + 21 * k + 9;
        C[k][0] = (t[0] * R[0] + t[1] * R[3]) + t[2] * R[6];
        C[k][1] = (t[0] * R[1] + t[1] * R[4]) + t[2] * R[7];
        C[k][2] = (t[0] * R[2] + t[1] * R[5]) + t[2] * R[8];
    }
    float B = 0.0f;
    for (int k = 0; k < ncal; k++) {
        float dx = C[0][0] - C[k][0], dy = C[0][1] - C[k][1], dz = C[0][2] - C[k][2];
        float d2 = dz * dz + (dx * dx + dy * dy);
        float d = lo_homography_sqrt(d2);
        if (!(d <= B)) B = d;
    }
    float x1_0 = lo_rcp_nr(near_z), x1_1 = lo_rcp_nr(far_z);
    float inv = x1_0 - x1_1;
    int n = f2i_sat(((inv * density) * B) * calibs[0]); /* the runtime's saturating (int) */
    if (n > 0x1000) n = 0x1000;
    if (n <= 0) return -3;
    n = ((n + align - 1) / align) * align;
    *n_out = n;
    if (!planes) return 0;
    if (cap < n) return -4;
    float step = (x1_0 - x1_1) / (float)(n - 1), bas = x1_1;
    if (n - 1 <= 7) {
        float v = bas;
        for (int i = 0; i < n - 1; i++) { planes[i] = 1.0f / v; v += step; }
    } else {
        int m = (n - 1) & ~7;
        float v4[4], d4 = 4.0f * step;
        for (int l = 0; l < 4; l++) v4[l] = step * (float)l + bas;
        for (int i = 0; i < m; i += 8) {
            float hi[4];
            for (int l = 0; l < 4; l++) hi[l] = v4[l] + d4;
            for (int l = 0; l < 4; l++) { planes[i + l] = lo_rcp_nr(v4[l]); planes[i + 4 + l] = lo_rcp_nr(hi[l]); }
            for (int l = 0; l < 4; l++) v4[l] = hi[l] + d4;
        }
        float v = (float)m * step + bas;
        for (int i = m; i < n - 1; i++) { planes[i] = 1.0f / v; v += step; }
    }
    planes[n - 1] = near_z;
    return 0;
}

static void lo_mul4(const double* a, const double* b, double* o) /* ViewTransform.cs:88-94 */
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            o[4 * i + j] = ((a[4 * i + 3] * b[12 + j] + a[4 * i + 2] * b[8 + j]) + (a[4 * i + 1] * b[4 + j] + a[4 * i] * b[j]));
}

static void lo_inverse4(const double* m, double* o) /* ViewTransform.cs:96-118 */
{
    double m00 = m[0], m01 = m[1], m02 = m[2], m03 = m[3], m10 = m[4], m11 = m[5], m12 = m[6], m13 = m[7], m20 = m[8], m21 = m[9],
           m22 = m[10], m23 = m[11], m30 = m[12], m31 = m[13], m32 = m[14], m33 = m[15];
    double adj[16];
    adj[0] = (m11 * (m22 * m33 - m23 * m32)) + ((m12 * (m23 * m31 - m33 * m21)) + (m13 * (m32 * m21 - m31 * m22)));
    adj[1] = (m01 * (m23 * m32 - m22 * m33)) + ((m02 * (m33 * m21 - m23 * m31)) + (m03 * (m31 * m22 - m32 * m21)));
    adj[2] = (((m02 * m13 - m12 * m03) * m31) + ((m12 * m01) * m33 + (m11 * m03) * m32)) - ((m01 * m13) * m32 + (m11 * m02) * m33);
    adj[3] = ((m23 * (m11 * m02 - m12 * m01)) + (m21 * (m12 * m03 - m02 * m13))) + (m22 * (m01 * m13 - m11 * m03));
    adj[4] = ((m12 * (m33 * m20 - m23 * m30)) + (m13 * (m22 * m30 - m32 * m20))) + (m10 * (m23 * m32 - m22 * m33));
    adj[5] = (m00 * (m22 * m33 - m23 * m32)) + ((m02 * (m23 * m30 - m33 * m20)) + (m03 * (m32 * m20 - m22 * m30)));
    adj[6] = (m33 * (m02 * m10 - m00 * m12)) + ((m30 * (m12 * m03 - m02 * m13)) + (m32 * (m00 * m13 - m03 * m10)));
    adj[7] = (((m02 * m13 - m12 * m03) * m20) + ((m03 * m10) * m22 + (m00 * m12) * m23)) - (m23 * (m02 * m10) + (m00 * m13) * m22);
    adj[8] = (m11 * (m23 * m30 - m33 * m20)) + ((m13 * (m31 * m20 - m21 * m30)) + (m10 * (m33 * m21 - m23 * m31)));
    adj[9] = (m00 * (m23 * m31 - m33 * m21)) + (((m33 * m20 - m23 * m30) * m01) + (m03 * (m21 * m30 - m31 * m20)));
    adj[10] = (((m01 * m13 - m11 * m03) * m30) + ((m00 * m11) * m33 + (m03 * m10) * m31)) - ((m00 * m13) * m31 + m33 * (m01 * m10));
    adj[11] = (m23 * (m01 * m10 - m00 * m11)) + ((m20 * (m11 * m03 - m01 * m13)) + (m21 * (m00 * m13 - m03 * m10)));
    adj[12] = (m11 * (m32 * m20 - m22 * m30)) + ((m12 * (m21 * m30 - m31 * m20)) + (m10 * (m31 * m22 - m32 * m21)));
    adj[13] = (m00 * (m32 * m21 - m31 * m22)) + (((m22 * m30 - m32 * m20) * m01) + (m02 * (m31 * m20 - m21 * m30)));
    adj[14] = (m32 * (m01 * m10 - m00 * m11)) + (((m11 * m02 - m12 * m01) * m30) + (m31 * (m00 * m12 - m02 * m10)));
    adj[15] = (((m12 * m01 - m11 * m02) * m20) + ((m02 * m10) * m21 + (m00 * m11) * m22)) - ((m01 * m10) * m22 + (m00 * m12) * m21);
    double det = ((adj[12] * m03 + m02 * adj[8]) + (adj[4] * m01 + m00 * adj[0]));
    double rd = 1.0 / det;
    for (int k = 0; k < 16; k++) o[k] = adj[k] * rd;
}

void lo_flow_matrix(const float* a, const float* b, float* out) /* ViewTransform.cs:120-133; records K[9], T[3], R[9] */
{
    double Ka[16] = {0}, Ea[16] = {0}, Kb[16] = {0}, Eb[16] = {0}, P1[16], P2[16], I2[16], M[16];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Ka[4 * i + j] = a[3 * i + j];
            Kb[4 * i + j] = b[3 * i + j];
            Ea[4 * i + j] = a[12 + 3 * i + j];
            Eb[4 * i + j] = b[12 + 3 * i + j];
        }
    for (int i = 0; i < 3; i++) { Ea[4 * i + 3] = a[9 + i]; Eb[4 * i + 3] = b[9 + i]; }
    Ka[15] = Kb[15] = Ea[15] = Eb[15] = 1.0;
    lo_mul4(Kb, Eb, P1);
    lo_mul4(Ka, Ea, P2);
    lo_inverse4(P2, I2);
    lo_mul4(P1, I2, M);
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) out[4 * c + r] = (float)M[4 * r + c];
}

uint64_t lo_approx_mismatches(int which, uint64_t lo, uint64_t hi) /* test helper: model vs the local instruction */
{
    uint64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t i = (int64_t)lo; i < (int64_t)hi; i++) {
        const uint32_t b = (uint32_t)i;
        const uint32_t want = which ? lo_rsqrtss_hw(b) : lo_rcpss_hw(b);
        const uint32_t got = which ? lo_rsqrt_bits(b) : lo_rcp_bits(b);
        bad += want != got;
    }
    return bad;
}
