// rcpss on the device: Intel's reciprocal approximation (IntelApprox.cs:20-28), from the 4096-entry table the
// instruction was read into (rcp_table.h) in constant memory, scaled by the input's power of two. Each translation
// unit that includes this holds its own copy of the table.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rcp_table.h"

namespace lux {
namespace {

__constant__ uint32_t kRcpDev[4096] = LUX_RCP_TABLE_INIT;

// +-0 and denormals -> +-inf, +-inf -> +-0, NaN -> quiet NaN, a result below the normal range -> +-0.
__device__ __forceinline__ float rcp_intel(float x)
{
    const uint32_t b = (uint32_t)__float_as_int(x);
    const uint32_t sign = b & 0x80000000u, e = (b >> 23) & 0xffu, m = b & 0x7fffffu;
    uint32_t r;
    if (e == 0xffu) r = m ? (b | 0x400000u) : sign;
    else if (e == 0) r = sign | 0x7f800000u;
    else {
        const uint32_t t = kRcpDev[m >> 11];
        const int re = (int)((t >> 23) & 0xffu) - ((int)e - 127);
        r = re <= 0 ? sign : (sign | ((uint32_t)re << 23) | (t & 0x7fffffu));
    }
    return __int_as_float((int)r);
}

// The same on the host, for the few values the host needs (plane boundaries).
inline float rcp_intel_host(float x)
{
    uint32_t b;
    __builtin_memcpy(&b, &x, 4);
    const uint32_t sign = b & 0x80000000u, e = (b >> 23) & 0xffu, m = b & 0x7fffffu;
    uint32_t r;
    if (e == 0xffu) r = m ? (b | 0x400000u) : sign;
    else if (e == 0) r = sign | 0x7f800000u;
    else {
        const uint32_t t = kRcpTable[m >> 11];
        const int re = (int)((t >> 23) & 0xffu) - ((int)e - 127);
        r = re <= 0 ? sign : (sign | ((uint32_t)re << 23) | (t & 0x7fffffu));
    }
    float f;
    __builtin_memcpy(&f, &r, 4);
    return f;
}

}  // namespace
}  // namespace lux
