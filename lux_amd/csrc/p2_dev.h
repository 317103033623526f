// The image-adaptive P2 of one path step (DenseStereoSgm.cs:267-274), shared by the path sweeps
// (kernels_impl.h) and the line-buffer-literal kernels (literal.hip).
#pragma once
#include "kernels.h"

namespace lux {
namespace {

// The adaptive P2 of one path step, written one rounded operation at a time like the reference (no FMA):
// E = 2^e by a cubic on the fraction plus an exponent add; note ei = (int)e - 1 for every negative e, so a
// negative whole e gets f = 1 (:269-272).
__device__ __forceinline__ uint32_t p2_step(const AdaptP2& q, uint32_t p2, uint32_t g, uint32_t gp)
{
    if (!q.on) return p2;
    const float d0 = __fmul_rn(__fsub_rn((float)(g & 255u), (float)(gp & 255u)), q.w0);
    const float d1 = __fmul_rn(__fsub_rn((float)((g >> 8) & 255u), (float)((gp >> 8) & 255u)), q.w1);
    const float d2 = __fmul_rn(__fsub_rn((float)((g >> 16) & 255u), (float)((gp >> 16) & 255u)), q.w2);
    const float S = __fadd_rn(__fadd_rn(fabsf(d0), fabsf(d2)), fabsf(d1));
    const float e = fmaxf(fminf(-S, 128.0f), -126.0f);
    const int ei = (int)e + (e < 0.0f ? -1 : 0);
    const float f = __fsub_rn(e, (float)ei);
    const float c3 = __int_as_float(0x3d9fcb52), c2 = __int_as_float(0x3e677e26), c1 = __int_as_float(0x3f322226),
                c0 = __int_as_float(0x3f7ffb19);
    const float poly = __fadd_rn(__fmul_rn(__fadd_rn(__fmul_rn(__fadd_rn(__fmul_rn(c3, f), c2), f), c1), f), c0);
    const float E = __int_as_float((int)((uint32_t)__float_as_int(poly) + ((uint32_t)ei << 23)));
    return (uint32_t)(int)__fmul_rn(q.base, E);
}

}  // namespace
}  // namespace lux
