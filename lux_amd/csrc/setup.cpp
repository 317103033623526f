// Host-side layer set-up of liblux_sgm.so: the skip mask of a layer (DenseLayer.SkipMask, DenseStereoSgm.cs:64-87)
// with its 32-bit Mersenne twister (FundamentalMatrixFilter.cs:7-30, the std::mt19937 the mask was defined with).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/lux_sgm.h"

namespace {

// MT19937 (Matsumoto & Nishimura 1998): 624-word state, initialised from one seed by the published recurrence
// x[i] = 1812433253 * (x[i-1] ^ (x[i-1] >> 30)) + i, regenerated in blocks of 624, tempered on output.
class Twister {
  public:
    explicit Twister(uint32_t seed)
    {
        s_[0] = seed;
        for (int i = 1; i < kN; i++) s_[i] = 1812433253u * (s_[i - 1] ^ (s_[i - 1] >> 30)) + (uint32_t)i;
        i_ = kN;
    }
    uint32_t next()
    {
        if (i_ >= kN) regenerate();
        uint32_t y = s_[i_++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }

  private:
    static constexpr int kN = 624, kM = 397;
    void regenerate()
    {
        for (int k = 0; k < kN; k++) {
            const uint32_t y = (s_[k] & 0x80000000u) | (s_[(k + 1) % kN] & 0x7fffffffu);
            s_[k] = s_[(k + kM) % kN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        i_ = 0;
    }
    uint32_t s_[kN];
    int i_;
};

}  // namespace

extern "C" {

int lux_sgm_skip_mask(int32_t mode, int32_t w, int32_t h, uint8_t* out)
{
    if (w <= 0 || h <= 0 || !out) return LUX_SGM_EINVAL;
    const size_t n = (size_t)w * h;
    if (mode == 0) {  // every pixel computed
        std::memset(out, 0, n);
        return LUX_SGM_OK;
    }
    if (mode != 2) return LUX_SGM_EUNSUPPORTED;  // "unrecognized sampling pattern" (:69)
    std::memset(out, 0xff, n);
    // 64x64 tiles (the last row / column of tiles takes the remainder when it is over half a tile); each tile draws
    // from a freshly seeded generator one computed pixel per 2x2 cell: dy then dx, each the low bit of a draw
    int nx = w / 64 + (64 < 2 * (w % 64) ? 1 : 0), ny = h / 64 + (64 < 2 * (h % 64) ? 1 : 0);
    if (nx < 1) nx = 1;
    if (ny < 1) ny = 1;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int x0 = 64 * i, y0 = 64 * j;
            const int x1 = std::min(w, x0 + 64 * (i == nx - 1 ? 2 : 1)), y1 = std::min(h, y0 + 64 * (j == ny - 1 ? 2 : 1));
            Twister mt(5489u);
            for (int y = y0; y < y1; y += 2)
                for (int x = x0; x < x1; x += 2) {
                    const int dy = (int)(mt.next() & 1u), dx = (int)(mt.next() & 1u);
                    const int yy = y + dy, xx = x + dx;
                    if (xx < w && yy < h) out[(size_t)yy * w + xx] = 0;
                }
        }
    return LUX_SGM_OK;
}

uint32_t lux_sgm_mt19937_draw(uint32_t seed, uint32_t n)
{
    Twister mt(seed);
    uint32_t v = 0;
    for (uint32_t k = 0; k < n; k++) v = mt.next();
    return v;
}

}  // extern "C"
