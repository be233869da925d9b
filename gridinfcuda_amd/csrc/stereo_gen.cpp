// Seeded synthetic stereo generator (project-owned; the reference has no
// generator -- its example builds costs from an image pair with
// min(|L - R(x-k)|, 20), examples/stereo/stereo.cc:28-34, which this follows
// except that out-of-range disparities cost the truncation value).
//
// Random stream: SplitMix64 (Steele, Lea & Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014), state = seed.
//   uniform integer in [lo, hi]: lo + next() % (hi - lo + 1)
//   noise byte:                  next() >> 56
// Draw order: W*H left-noise bytes (row-major), then per rectangle
// x0, y0, width, height, disparity, then one noise byte per right-image hole
// (row-major).
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "ginf_bp.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    int uniform(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
    uint8_t byte() { return (uint8_t)(next() >> 56); }
};

}  // namespace

extern "C" int ginf_gen_stereo(uint64_t seed, int W, int H, int L, int num_rects, int dmax, int trunc, uint8_t *left,
                               uint8_t *right, int32_t *disp, float *cost_f32, int32_t *cost_i32) {
    if (W < 1 || H < 1 || L < 1 || num_rects < 0 || dmax < 0) return GINF_ERR_INVALID;
    SplitMix64 rng{seed};
    const size_t n = (size_t)W * H;
    std::vector<uint8_t> raw(n), lft(n), rgt(n);
    std::vector<int32_t> dsp(n);
    std::vector<uint8_t> filled(n, 0);
    for (size_t i = 0; i < n; ++i) raw[i] = rng.byte();
    // 3x3 box filter, clamp-to-edge, rounded: (sum + 4) / 9.
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int acc = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int yy = y + dy < 0 ? 0 : (y + dy >= H ? H - 1 : y + dy);
                    int xx = x + dx < 0 ? 0 : (x + dx >= W ? W - 1 : x + dx);
                    acc += raw[(size_t)yy * W + xx];
                }
            lft[(size_t)y * W + x] = (uint8_t)((acc + 4) / 9);
        }
    const int background = dmax < 2 ? dmax : 2;
    for (size_t i = 0; i < n; ++i) dsp[i] = background;
    const int wmin = W / 16 > 1 ? W / 16 : 1, wmax = W / 4 > wmin ? W / 4 : wmin;
    const int hmin = H / 16 > 1 ? H / 16 : 1, hmax = H / 4 > hmin ? H / 4 : hmin;
    for (int r = 0; r < num_rects; ++r) {
        const int x0 = rng.uniform(0, W - 1), y0 = rng.uniform(0, H - 1);
        const int rw = rng.uniform(wmin, wmax), rh = rng.uniform(hmin, hmax);
        const int dv = rng.uniform(0, dmax);
        for (int y = y0; y < y0 + rh && y < H; ++y)
            for (int x = x0; x < x0 + rw && x < W; ++x) dsp[(size_t)y * W + x] = dv;
    }
    // Forward warp R(x - d, y) = L(x, y); painting in increasing d lets the
    // nearer (larger-d) pixel win, i.e. each target keeps its largest-d source.
    std::vector<int32_t> best(n, -1);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int dv = dsp[(size_t)y * W + x];
            const int xt = x - dv;
            if (xt < 0) continue;
            const size_t t = (size_t)y * W + xt;
            if (dv > best[t]) { best[t] = dv; rgt[t] = lft[(size_t)y * W + x]; filled[t] = 1; }
        }
    for (size_t i = 0; i < n; ++i)
        if (!filled[i]) rgt[i] = rng.byte();
    if (cost_f32 || cost_i32) {
        for (int l = 0; l < L; ++l)
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    int c = trunc;
                    if (x - l >= 0) {
                        c = abs((int)lft[(size_t)y * W + x] - (int)rgt[(size_t)y * W + x - l]);
                        if (c > trunc) c = trunc;
                    }
                    const size_t o = ((size_t)l * H + y) * W + x;
                    if (cost_f32) cost_f32[o] = (float)c;
                    if (cost_i32) cost_i32[o] = c;
                }
    }
    for (size_t i = 0; i < n; ++i) {
        if (left) left[i] = lft[i];
        if (right) right[i] = rgt[i];
        if (disp) disp[i] = dsp[i];
    }
    return GINF_OK;
}
