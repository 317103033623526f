// Host arithmetic of the L16 layer set-up that feeds the plane-sweep cost: the plane list (DenseStereo.Planes,
// DenseStereo.cs:41-88), the reference-to-camera warps (Mat4D.FlowMatrix, ViewTransform.cs:86-133), the baseline's
// square root (Homography.Sqrt, Homography.cs:18) and the rcpss / rsqrtss models they use (IntelApprox.RcpBits /
// RsqrtBits, IntelApprox.cs:21-44). Computed once per capture, on the host as in the reference; every float operation
// is one rounded operation in the reference's order (this file is compiled with -ffp-contract=off: no FMA).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/lux_sgm.h"
#include "rcp_table.h"
#include "rsqrt_table.h"

namespace {

inline float f_of(uint32_t b)
{
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
inline uint32_t b_of(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

// rcpss: the table entry of the top 12 mantissa bits scaled by the input's power of two; +-0 and denormals -> +-inf,
// +-inf -> +-0, NaN -> the quieted NaN, a denormal result -> +-0 (IntelApprox.cs:21-29).
uint32_t rcp_bits(uint32_t b)
{
    const uint32_t sign = b & 0x80000000u, e = (b >> 23) & 0xffu, m = b & 0x7fffffu;
    if (e == 0xffu) return m ? (b | 0x400000u) : sign;
    if (e == 0) return sign | 0x7f800000u;
    const uint32_t r = kRcpTable[m >> 11];
    const int re = (int)((r >> 23) & 0xffu) - ((int)e - 127);
    if (re <= 0) return sign;
    return sign | ((uint32_t)re << 23) | (r & 0x7fffffu);
}

// rsqrtss: one table per exponent parity; NaN -> quiet NaN, +-0 and denormals -> +-inf, a negative input -> the
// indefinite QNaN, +inf -> +0, a denormal result -> +0 (IntelApprox.cs:32-44).
uint32_t rsqrt_bits(uint32_t b)
{
    const uint32_t sign = b & 0x80000000u, e = (b >> 23) & 0xffu, m = b & 0x7fffffu;
    if (e == 0xffu && m) return b | 0x400000u;
    if (e == 0) return sign | 0x7f800000u;
    if (sign) return 0xffc00000u;
    if (e == 0xffu) return 0;
    const int k = (int)e - 127;
    uint32_t r;
    int re;
    if ((k & 1) == 0) {
        r = kRsqrtOddTable[m >> 11];
        re = (int)((r >> 23) & 0xffu) - k / 2;
    } else {
        r = kRsqrtEvenTable[m >> 11];
        re = (int)((r >> 23) & 0xffu) - (k - 1) / 2;
    }
    if (re <= 0) return 0;
    return ((uint32_t)re << 23) | (r & 0x7fffffu);
}

inline float rcp(float x) { return f_of(rcp_bits(b_of(x))); }

// One Newton step on rcpss: x0 + x0 * (1 - a * x0) (DenseStereo.RcpNr, DenseStereo.cs:36-37).
inline float rcp_nr(float a)
{
    const float x0 = rcp(a);
    return x0 + x0 * (1.0f - a * x0);
}

// Homography.Sqrt (Homography.cs:18): x * rsqrt(x) refined by one Newton step, 0 for x == 0.
float homography_sqrt(float x)
{
    const float r = f_of(rsqrt_bits(b_of(x)));
    const float A = x * r, B = A * (-0.5f), C = A * r, D = C + (-3.0f), s = D * B;
    return x == 0.0f ? 0.0f : s;
}

// float -> int as the reference's runtime converts (truncation toward zero, saturating at the int range, NaN -> 0):
// an overflowing plane count becomes the 4096 cap, a NaN one the "baseline too small" error.
int to_int_sat(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f < -2147483648.0f) return INT32_MIN;
    return (int)f;
}

// Mat4D (ViewTransform.cs:86-118): 4x4 double products as ((k3 + k2) + (k1 + k0)) and the adjugate inverse, each
// expression in the reference's order.
void mul4(const double* a, const double* b, double* o)
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            o[4 * i + j] = ((a[4 * i + 3] * b[12 + j] + a[4 * i + 2] * b[8 + j]) + (a[4 * i + 1] * b[4 + j] + a[4 * i] * b[j]));
}

void inverse4(const double* m, double* o)
{
    const double m00 = m[0], m01 = m[1], m02 = m[2], m03 = m[3], m10 = m[4], m11 = m[5], m12 = m[6], m13 = m[7];
    const double m20 = m[8], m21 = m[9], m22 = m[10], m23 = m[11], m30 = m[12], m31 = m[13], m32 = m[14], m33 = m[15];
    double a[16];
    a[0] = (m11 * (m22 * m33 - m23 * m32)) + ((m12 * (m23 * m31 - m33 * m21)) + (m13 * (m32 * m21 - m31 * m22)));
    a[1] = (m01 * (m23 * m32 - m22 * m33)) + ((m02 * (m33 * m21 - m23 * m31)) + (m03 * (m31 * m22 - m32 * m21)));
    a[2] = (((m02 * m13 - m12 * m03) * m31) + ((m12 * m01) * m33 + (m11 * m03) * m32)) - ((m01 * m13) * m32 + (m11 * m02) * m33);
    a[3] = ((m23 * (m11 * m02 - m12 * m01)) + (m21 * (m12 * m03 - m02 * m13))) + (m22 * (m01 * m13 - m11 * m03));
    a[4] = ((m12 * (m33 * m20 - m23 * m30)) + (m13 * (m22 * m30 - m32 * m20))) + (m10 * (m23 * m32 - m22 * m33));
    a[5] = (m00 * (m22 * m33 - m23 * m32)) + ((m02 * (m23 * m30 - m33 * m20)) + (m03 * (m32 * m20 - m22 * m30)));
    a[6] = (m33 * (m02 * m10 - m00 * m12)) + ((m30 * (m12 * m03 - m02 * m13)) + (m32 * (m00 * m13 - m03 * m10)));
    a[7] = (((m02 * m13 - m12 * m03) * m20) + ((m03 * m10) * m22 + (m00 * m12) * m23)) - (m23 * (m02 * m10) + (m00 * m13) * m22);
    a[8] = (m11 * (m23 * m30 - m33 * m20)) + ((m13 * (m31 * m20 - m21 * m30)) + (m10 * (m33 * m21 - m23 * m31)));
    a[9] = (m00 * (m23 * m31 - m33 * m21)) + (((m33 * m20 - m23 * m30) * m01) + (m03 * (m21 * m30 - m31 * m20)));
    a[10] = (((m01 * m13 - m11 * m03) * m30) + ((m00 * m11) * m33 + (m03 * m10) * m31)) - ((m00 * m13) * m31 + m33 * (m01 * m10));
    a[11] = (m23 * (m01 * m10 - m00 * m11)) + ((m20 * (m11 * m03 - m01 * m13)) + (m21 * (m00 * m13 - m03 * m10)));
    a[12] = (m11 * (m32 * m20 - m22 * m30)) + ((m12 * (m21 * m30 - m31 * m20)) + (m10 * (m31 * m22 - m32 * m21)));
    a[13] = (m00 * (m32 * m21 - m31 * m22)) + (((m22 * m30 - m32 * m20) * m01) + (m02 * (m31 * m20 - m21 * m30)));
    a[14] = (m32 * (m01 * m10 - m00 * m11)) + (((m11 * m02 - m12 * m01) * m30) + (m31 * (m00 * m12 - m02 * m10)));
    a[15] = (((m12 * m01 - m11 * m02) * m20) + ((m02 * m10) * m21 + (m00 * m11) * m22)) - ((m01 * m10) * m22 + (m00 * m12) * m21);
    const double det = ((a[12] * m03 + m02 * a[8]) + (a[4] * m01 + m00 * a[0]));
    const double rd = 1.0 / det;
    for (int k = 0; k < 16; k++) o[k] = a[k] * rd;
}

// K (3x3) and [R | T] as 4x4 doubles (ViewTransform.cs:122-123).
void kd(const lux_sgm_calib& c, double* k)
{
    for (int i = 0; i < 16; i++) k[i] = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) k[4 * i + j] = c.K[3 * i + j];
    k[15] = 1.0;
}
void ed(const lux_sgm_calib& c, double* e)
{
    for (int i = 0; i < 16; i++) e[i] = 0.0;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) e[4 * i + j] = c.R[3 * i + j];
        e[4 * i + 3] = c.T[i];
    }
    e[15] = 1.0;
}

thread_local std::string g_err;

}  // namespace

uint32_t lux_sgm_rcpss_bits(uint32_t b) { return rcp_bits(b); }
uint32_t lux_sgm_rsqrtss_bits(uint32_t b) { return rsqrt_bits(b); }
float lux_sgm_homography_sqrt(float x) { return homography_sqrt(x); }

int lux_sgm_flow_matrix(const lux_sgm_calib* a, const lux_sgm_calib* b, float* out)
{
    if (!a || !b || !out) return LUX_SGM_EINVAL;
    double kb[16], eb[16], ka[16], ea[16], p1[16], p2[16], inv[16], m[16];
    kd(*b, kb);
    ed(*b, eb);
    kd(*a, ka);
    ed(*a, ea);
    mul4(kb, eb, p1);
    mul4(ka, ea, p2);
    inverse4(p2, inv);
    mul4(p1, inv, m);
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) out[4 * c + r] = (float)m[4 * r + c];  // column-major float
    return LUX_SGM_OK;
}

int lux_sgm_planes(float near_z, float far_z, const lux_sgm_calib* calibs, int32_t n_calibs, float density, int32_t align,
                   float* planes, int32_t cap, int32_t* n_out)
{
    if (!calibs || n_calibs <= 0 || !n_out || align <= 0 || align > 4096) {
        g_err = "bad arguments";
        return LUX_SGM_EINVAL;
    }
    if (!(near_z < far_z)) {
        g_err = "The near end has to be smaller than the fard end!";
        return LUX_SGM_EINVAL;
    }
    if (!(near_z > 0.0f && far_z > 0.0f)) {
        g_err = "Both the near and far end have to be positive!";
        return LUX_SGM_EINVAL;
    }
    // camera centres C = -R^T t up to sign, as the reference forms them (:47-52), and the largest distance to camera 0
    float B = 0.0f, c0[3] = {0, 0, 0};
    for (int k = 0; k < n_calibs; k++) {
        const float *R = calibs[k].R, *t = calibs[k].T;
        const float c[3] = {(t[0] * R[0] + t[1] * R[3]) + t[2] * R[6], (t[0] * R[1] + t[1] * R[4]) + t[2] * R[7],
                            (t[0] * R[2] + t[1] * R[5]) + t[2] * R[8]};
        if (k == 0) std::memcpy(c0, c, sizeof c);
        const float dx = c0[0] - c[0], dy = c0[1] - c[1], dz = c0[2] - c[2];
        const float d2 = dz * dz + (dx * dx + dy * dy);
        const float d = homography_sqrt(d2);
        if (!(d <= B)) B = d;  // maxss order: a NaN distance is taken
    }
    const float xn = rcp_nr(near_z), xf = rcp_nr(far_z);
    const float inv = xn - xf;
    int n = to_int_sat(((inv * density) * B) * calibs[0].K[0]);
    if (n > 0x1000) n = 0x1000;
    if (n <= 0) {
        g_err = "Baseline too small, check calibration.";
        return LUX_SGM_EINVAL;
    }
    n = ((n + align - 1) / align) * align;
    *n_out = n;
    if (!planes) return LUX_SGM_OK;  // a size query
    if (cap < n) {
        g_err = "plane array too small";
        return LUX_SGM_EINVAL;
    }
    const float step = (xn - xf) / (float)(n - 1), bas = xf;
    if (n - 1 <= 7) {
        float v = bas;
        for (int i = 0; i < n - 1; i++) {
            planes[i] = 1.0f / v;
            v += step;
        }
    } else {
        // eight planes a step: two four-lane groups of 1/z, each refined from rcpps (:66-75)
        const int m = (n - 1) & ~7;
        float v4[4];
        for (int l = 0; l < 4; l++) v4[l] = step * (float)l + bas;
        const float d4 = 4.0f * step;
        for (int i = 0; i < m; i += 8) {
            float hi[4];
            for (int l = 0; l < 4; l++) hi[l] = v4[l] + d4;
            for (int l = 0; l < 4; l++) {
                planes[i + l] = rcp_nr(v4[l]);
                planes[i + 4 + l] = rcp_nr(hi[l]);
            }
            for (int l = 0; l < 4; l++) v4[l] = hi[l] + d4;
        }
        float v = (float)m * step + bas;
        for (int i = m; i < n - 1; i++) {
            planes[i] = 1.0f / v;
            v += step;
        }
    }
    planes[n - 1] = near_z;
    return LUX_SGM_OK;
}

const char* lux_sgm_calib_last_error(void) { return g_err.c_str(); }
