// BF16 3x3 convolution kernels on MFMA (gfx950 v_mfma_f32_16x16x32_bf16,
// FP32 accumulate) for the learned warm start, plus the small elementwise
// kernels around them (pooling, ReLU backward, loss, Adam, weight
// preparation, message injection into the BP state).  No reference
// counterpart: the warm start is a project-owned component.
//
// Tensors are NHWC (channels fastest); weights are [Cout][3][3][Cin].
// MFMA operand layout (16x16x32): lane l holds A[l%16][8*(l/16) .. +7] and
// B[8*(l/16) .. +7][l%16]; the result lane l holds D[4*(l/16) + i][l%16],
// i = 0..3.  A rows are pixels, B columns output channels, K runs over
// (kernel tap, input channel) in steps of 32 channels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "ginf_bp.h"
#include "bp_layout.h"
#include "conv_common.h"

namespace {

using ginf::bf16x8;
using ginf::f4;
using ginf::bf2f;
using ginf::f2bf;
using ginf::load8;

}  // namespace

namespace ginf {
// convnet_fused.hip
hipError_t launch_convnet_fused(int layers, const int *chans, const uint16_t *const *w, const float *const *b,
                                const uint16_t *x, int N, int H, int W, float *out, hipStream_t st);
}  // namespace ginf

namespace {

// y = conv3x3(x, w) + bias (zero padding 1), optional ReLU.  Each wave owns
// 16 consecutive pixels (flattened n,h,w) and 32 output channels (two MFMA
// column tiles); blockIdx.y selects the 32-channel group, so a wide output
// layer spreads over more waves.
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
                                                 const float *__restrict__ bias, int N, int H, int W, int Cin,
                                                 int Cout, int relu, uint16_t *__restrict__ yb,
                                                 float *__restrict__ yf) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t P = (int64_t)N * H * W;
    const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (p0 >= P) return;
    const int cg = blockIdx.y;  // output channels 32*cg .. 32*cg+31
    const int m = lane & 15, kg = lane >> 4;
    const int64_t p = p0 + m;
    const bool valid = p < P;
    const int n = valid ? (int)(p / ((int64_t)H * W)) : 0;
    const int rem = valid ? (int)(p - (int64_t)n * H * W) : 0;
    const int h = rem / W, wc = rem - (rem / W) * W;
    f4 acc0 = f4{0.f, 0.f, 0.f, 0.f}, acc1 = f4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero = __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
    const int64_t wstride = (int64_t)16 * 9 * Cin;  // one 16-channel tile of weights
    for (int kh = 0; kh < 3; ++kh) {
        for (int kw = 0; kw < 3; ++kw) {
            const int hh = h + kh - 1, ww = wc + kw - 1;
            const bool inb = valid && hh >= 0 && hh < H && ww >= 0 && ww < W;
            const uint16_t *xp = x + (((int64_t)n * H + (inb ? hh : 0)) * W + (inb ? ww : 0)) * Cin + kg * 8;
            const uint16_t *wp = w + ((int64_t)(cg * 32 + m) * 9 + kh * 3 + kw) * Cin + kg * 8;
            for (int ci0 = 0; ci0 < Cin; ci0 += 32) {
                const bf16x8 a = inb ? load8(xp + ci0) : zero;
                const bf16x8 b0 = load8(wp + ci0);
                const bf16x8 b1 = load8(wp + wstride + ci0);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int co = cg * 32 + t * 16 + m;
        const float bv = bias ? bias[co] : 0.f;
        const f4 acc = t ? acc1 : acc0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t pp = p0 + kg * 4 + i;
            if (pp >= P) continue;
            float v = acc[i] + bv;
            if (relu) v = v > 0.f ? v : 0.f;
            if (yb) yb[pp * Cout + co] = f2bf(v);
            if (yf) yf[pp * Cout + co] = v;
        }
    }
}

// The same convolution with every operand of a group of taps loaded before
// the group's MFMAs: the loop above waits for one memory round trip per
// (tap, 32-channel step), which bounds a layer at 1/4 resolution (one wave
// per SIMD, nothing to hide the latency with) to about 13 us.  CS = Cin / 32
// channel steps; TAPS = taps per group (all nine for CS = 1, one kernel row
// otherwise, so the operands fit the register file).  The MFMAs run in the
// same order as above (tap-major, channel steps inside), so the result is
// identical.
template <int CS, int TAPS>
__global__ __launch_bounds__(256) void k_conv3x3_pf(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
                                                    const float *__restrict__ bias, int N, int H, int W, int Cout,
                                                    int relu, uint16_t *__restrict__ yb, float *__restrict__ yf) {
    constexpr int Cin = 32 * CS;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t P = (int64_t)N * H * W;
    const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (p0 >= P) return;
    const int cg = blockIdx.y;
    const int m = lane & 15, kg = lane >> 4;
    const int64_t p = p0 + m;
    const bool valid = p < P;
    const int n = valid ? (int)(p / ((int64_t)H * W)) : 0;
    const int rem = valid ? (int)(p - (int64_t)n * H * W) : 0;
    const int h = rem / W, wc = rem - (rem / W) * W;
    f4 acc0 = f4{0.f, 0.f, 0.f, 0.f}, acc1 = f4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero = __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
    const int64_t wstride = (int64_t)16 * 9 * Cin;
#pragma unroll
    for (int g0 = 0; g0 < 9; g0 += TAPS) {
        bf16x8 a[TAPS][CS], b0[TAPS][CS], b1[TAPS][CS];
#pragma unroll
        for (int t = 0; t < TAPS; ++t) {
            const int kh = (g0 + t) / 3, kw = (g0 + t) % 3;
            const int hh = h + kh - 1, ww = wc + kw - 1;
            const bool inb = valid && hh >= 0 && hh < H && ww >= 0 && ww < W;
            // out-of-image taps read a clamped, in-bounds address and are replaced by zero
            const uint16_t *xp = x + (((int64_t)n * H + (inb ? hh : 0)) * W + (inb ? ww : 0)) * Cin + kg * 8;
            const uint16_t *wp = w + ((int64_t)(cg * 32 + m) * 9 + kh * 3 + kw) * Cin + kg * 8;
#pragma unroll
            for (int c = 0; c < CS; ++c) {
                a[t][c] = inb ? load8(xp + 32 * c) : zero;
                b0[t][c] = load8(wp + 32 * c);
                b1[t][c] = load8(wp + wstride + 32 * c);
            }
        }
#pragma unroll
        for (int t = 0; t < TAPS; ++t)
#pragma unroll
            for (int c = 0; c < CS; ++c) {
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t][c], b0[t][c], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t][c], b1[t][c], acc1, 0, 0, 0);
            }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int co = cg * 32 + t * 16 + m;
        const float bv = bias ? bias[co] : 0.f;
        const f4 acc = t ? acc1 : acc0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t pp = p0 + kg * 4 + i;
            if (pp >= P) continue;
            float v = acc[i] + bv;
            if (relu) v = v > 0.f ? v : 0.f;
            if (yb) yb[pp * Cout + co] = f2bf(v);
            if (yf) yf[pp * Cout + co] = v;
        }
    }
}

// dW[co][kh][kw][ci] = sum_p dz[p][co] * x[p + (kh-1, kw-1)][ci]: one block per
// (16-channel co tile, 16-channel ci tile, tap); its four waves split the
// pixels and add their partial sums in LDS (fixed order, deterministic).
// The pixels are split into gridDim.y chunks of `chunk` pixels: block
// (tile, chunk) writes its chunk's sum to part[chunk][co][tap][ci] (or
// straight to dw with one chunk), and k_dw_reduce adds the chunks in index
// order, so the result does not depend on scheduling.  A layer at 160x120 x
// 16 spread over (co, ci, tap) tiles alone is 36-72 blocks each walking all
// 307,200 pixels (105.6 us per layer).
__global__ __launch_bounds__(256) void k_conv3x3_dw(const uint16_t *__restrict__ x, const uint16_t *__restrict__ dz,
                                                    int N, int H, int W, int Cin, int Cout, int64_t chunk,
                                                    float *__restrict__ dw) {
    __shared__ float red[4][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ct = Cout / 16, it = Cin / 16;
    int b = blockIdx.x;
    const int tap = b % 9; b /= 9;
    const int ti = b % it; b /= it;
    const int to = b;
    const int kh = tap / 3, kw = tap % 3;
    const int m = lane & 15, kg = lane >> 4;
    const int64_t HW = (int64_t)H * W;
    const int64_t k0 = (int64_t)blockIdx.y * chunk;
    const int64_t P = k0 + chunk < (int64_t)N * HW ? k0 + chunk : (int64_t)N * HW;  // end of this chunk
    f4 acc = f4{0.f, 0.f, 0.f, 0.f};
    for (int64_t q0 = k0 + (int64_t)wave * 32; q0 < P; q0 += 4 * 32) {
        uint16_t av[8], bv[8];
        // row and column of this lane's first pixel, then stepped (no division per pixel)
        const int64_t pb = q0 + kg * 8;
        const int rem0 = (int)(pb % HW);
        int h0 = rem0 / W, w0 = rem0 - (rem0 / W) * W;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t p = pb + j;
            uint16_t a = 0, bb = 0;
            if (p < P) {
                a = dz[p * Cout + to * 16 + m];
                const int hh = h0 + kh - 1, ww = w0 + kw - 1;
                if (hh >= 0 && hh < H && ww >= 0 && ww < W)
                    bb = x[(p + (int64_t)(kh - 1) * W + (kw - 1)) * Cin + ti * 16 + m];
            }
            av[j] = a;
            bv[j] = bb;
            if (++w0 == W) { w0 = 0; if (++h0 == H) h0 = 0; }
        }
        bf16x8 A, B;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            A[j] = __builtin_bit_cast(__bf16, av[j]);
            B[j] = __builtin_bit_cast(__bf16, bv[j]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][(kg * 4 + i) * 16 + m] = acc[i];
    __syncthreads();
    if (threadIdx.x < 256) {
        const int r = threadIdx.x / 16, c = threadIdx.x % 16;  // r: co within tile, c: ci within tile
        const float v = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
        const int co = to * 16 + r, ci = ti * 16 + c;
        dw[(int64_t)blockIdx.y * Cout * 9 * Cin + ((int64_t)co * 9 + tap) * Cin + ci] = v;
    }
}

// The weight gradient with its operands staged through LDS.  The MFMA
// reduces over K = pixels, so each lane needs eight consecutive pixels of one
// channel, while NHWC memory holds one pixel's channels contiguously.  Block
// (pixel block, co group of 16*CT channels) walks image patches of 4 rows x
// 32 columns: it loads the patch's dz and the x window around it (6 rows x 34
// columns x Cin) with 16-byte reads and stores both as they lie in memory,
// one pixel per LDS row, with one 16-byte LDS write per load.  The operands
// are read back transposed by ds_read_b64_tr_b16: in each group of 16 lanes,
// lane 4r + q names 4 channels of pixel r (r = 0..3) and lane j receives
// channel j of those 4 pixels, so two reads give a lane its 8 pixels of one
// channel.  A kernel tap is only an address offset into the x window.  The
// (tap, ci tile) pairs are dealt round-robin to the four waves and each wave
// does them for all CT co tiles, so every staged operand feeds several
// MFMAs.  Pixels outside the image are zero in both operands (the zero
// padding of the forward pass).  The block's sums go to
// part[block][co][tap][ci] (straight to dw with one block), which
// k_chunk_reduce adds in block order.  CIT = Cin / 16.
constexpr int kDwPR = 4, kDwPC = 32;  // patch rows, columns
constexpr int kDwBlocks = 256;          // pixel blocks x co groups per launch (about one per CU)

// LDS pitch (elements) of a pixel row of ch channels: the 16 lanes of a
// transposed read touch 4 rows x 32 bytes, which fall in distinct banks when
// the pitch in dwords is 8 or 24 modulo 32.
__host__ __device__ constexpr int dw_pitch(int ch) { return (ch / 2) % 16 == 0 ? ch + 16 : ch; }

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s16x4 *lds_s16x4;

// two transposed reads: channel (lane % 16) of pixels p .. p+3 and p+4 .. p+7,
// where p's row sits at `row` and the next pixels follow at `pitch` elements
__device__ __forceinline__ bf16x8 tr_read8(const uint16_t *row, int pitch) {
    const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4)row);
    const s16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4)(row + 4 * pitch));
    return __builtin_bit_cast(bf16x8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}

template <int CIT, int CT>
__global__ __launch_bounds__(256) void k_conv3x3_dw_tr(const uint16_t *__restrict__ x, const uint16_t *__restrict__ dz,
                                                       int N, int H, int W, int Cout, float *__restrict__ out) {
    constexpr int Cin = 16 * CIT, CO = 16 * CT;
    constexpr int WC = kDwPC + 2, WR = kDwPR + 2, C8 = Cin / 8, O8 = CO / 8;
    constexpr int XP = dw_pitch(Cin), AP = dw_pitch(CO);
    constexpr int NX = (WR * WC * C8 + 255) / 256;          // x window loads per thread
    constexpr int NA = (kDwPR * kDwPC * O8 + 255) / 256;    // dz loads per thread
    constexpr int NQ = 9 * CIT, QW = (NQ + 3) / 4;          // (tap, ci tile) pairs: all, per wave
    __shared__ __attribute__((aligned(16))) uint16_t sA[kDwPR * kDwPC * AP];
    __shared__ __attribute__((aligned(16))) uint16_t sX[WR * WC * XP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, rq = (lane >> 2) & 3, qq = lane & 3;
    const int co0 = blockIdx.y * CO;
    const int ppr = (W + kDwPC - 1) / kDwPC, prow = (H + kDwPR - 1) / kDwPR;
    const int64_t npatch = (int64_t)N * prow * ppr;
    f4 acc[CT][QW];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int t = 0; t < QW; ++t) acc[c][t] = f4{0.f, 0.f, 0.f, 0.f};
    // A patch's operands are read from HBM into registers one patch ahead of
    // the MFMAs that use them, so the loads of patch i+1 are in flight while
    // patch i is reduced.
    uint4 va[NA], vx[NX];
    auto fetch = [&](int64_t pi) {
        const int n = (int)(pi / ((int64_t)prow * ppr));
        const int rem = (int)(pi - (int64_t)n * prow * ppr);
        const int r0 = (rem / ppr) * kDwPR, c0 = (rem % ppr) * kDwPC;
#pragma unroll
        for (int u = 0; u < NA; ++u) {  // dz patch: 128 pixels x CO channels
            const int t = threadIdx.x + 256 * u, px = t / O8, h8 = t % O8;
            const int hh = r0 + px / kDwPC, ww = c0 + px % kDwPC;
            va[u] = uint4{0u, 0u, 0u, 0u};
            if (t < kDwPR * kDwPC * O8 && hh < H && ww < W)
                va[u] = *reinterpret_cast<const uint4 *>(dz + (((int64_t)n * H + hh) * W + ww) * Cout + co0 + 8 * h8);
        }
#pragma unroll
        for (int u = 0; u < NX; ++u) {  // x window: rows r0-1 .. r0+4, columns c0-1 .. c0+32
            const int t = threadIdx.x + 256 * u;
            const int ci8 = t % C8, wcr = t / C8;
            const int wc = wcr % WC, rr = wcr / WC;
            const int hh = r0 - 1 + rr, ww = c0 - 1 + wc;
            vx[u] = uint4{0u, 0u, 0u, 0u};
            if (t < WR * WC * C8 && hh >= 0 && hh < H && ww >= 0 && ww < W)
                vx[u] = *reinterpret_cast<const uint4 *>(x + (((int64_t)n * H + hh) * W + ww) * Cin + 8 * ci8);
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int t = threadIdx.x + 256 * u;
            if (t < kDwPR * kDwPC * O8) *reinterpret_cast<uint4 *>(&sA[(t / O8) * AP + 8 * (t % O8)]) = va[u];
        }
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int t = threadIdx.x + 256 * u;
            if (t < WR * WC * C8) *reinterpret_cast<uint4 *>(&sX[(t / C8) * XP + 8 * (t % C8)]) = vx[u];
        }
    };
    int64_t pi = blockIdx.x;
    if (pi < npatch) fetch(pi);
    for (; pi < npatch; pi += gridDim.x) {
        stage();
        __syncthreads();
        if (pi + gridDim.x < npatch) fetch(pi + gridDim.x);
#pragma unroll
        for (int r = 0; r < kDwPR; ++r) {
            bf16x8 A[CT];
#pragma unroll
            for (int c = 0; c < CT; ++c) A[c] = tr_read8(&sA[(r * kDwPC + 8 * g + rq) * AP + 16 * c + 4 * qq], AP);
#pragma unroll
            for (int t = 0; t < QW; ++t) {
                const int q = wave + 4 * t;
                if (q >= NQ) break;
                const int tap = q / CIT, ct = q % CIT, kh = tap / 3, kw = tap % 3;
                const bf16x8 B = tr_read8(&sX[((r + kh) * WC + 8 * g + rq + kw) * XP + 16 * ct + 4 * qq], XP);
#pragma unroll
                for (int c = 0; c < CT; ++c) acc[c][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[c], B, acc[c][t], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int64_t nout = (int64_t)Cout * 9 * Cin;
#pragma unroll
    for (int t = 0; t < QW; ++t) {
        const int q = wave + 4 * t;
        if (q >= NQ) break;
        const int tap = q / CIT, ct = q % CIT;
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int co = co0 + 16 * c + 4 * g + i;
                out[(int64_t)blockIdx.x * nout + ((int64_t)co * 9 + tap) * Cin + 16 * ct + (lane & 15)] = acc[c][t][i];
            }
    }
}

// out[o] = sum over chunks c = 0, 1, ... of part[c][o], in that order.
__global__ void k_chunk_reduce(const float *__restrict__ part, int nchunk, int64_t n, float *__restrict__ out) {
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (int64_t)gridDim.x * blockDim.x) {
        float v = 0.f;
        int c = 0;
        for (; c + 8 <= nchunk; c += 8) {  // eight loads in flight, added in order
            float t[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) t[k] = part[(int64_t)(c + k) * n + o];
#pragma unroll
            for (int k = 0; k < 8; ++k) v += t[k];
        }
        for (; c < nchunk; ++c) v += part[(int64_t)c * n + o];
        out[o] = v;
    }
}

// Bias gradient partials: block (channel group, pixel chunk) sums dz over its
// chunk for CW consecutive channels, 256 / CW pixel rows at a time (each row
// read as CW * 2 contiguous bytes); the rows' partial sums are added in a
// fixed order.  part[chunk][co].
__global__ __launch_bounds__(256) void k_bias_grad_part(const uint16_t *__restrict__ dz, int64_t P, int Cout, int CW,
                                                        int64_t chunk, float *__restrict__ part) {
    __shared__ float red[256];
    const int rows = 256 / CW;
    const int c = threadIdx.x % CW, r = threadIdx.x / CW;
    const int co = blockIdx.x * CW + c;
    const int64_t k0 = (int64_t)blockIdx.y * chunk;
    const int64_t k1 = k0 + chunk < P ? k0 + chunk : P;
    float acc = 0.f;
    for (int64_t p = k0 + r; p < k1; p += rows) acc += bf2f(dz[p * Cout + co]);
    red[threadIdx.x] = acc;
    __syncthreads();
    if (r == 0) {
        float v = red[c];
        for (int k = 1; k < rows; ++k) v += red[k * CW + c];
        part[(int64_t)blockIdx.y * Cout + co] = v;
    }
}


// dz = bf16(dy * (y > 0)) for a ReLU layer (y its output), bf16(dy) otherwise.
__global__ void k_relu_bwd(const float *__restrict__ dy, const uint16_t *__restrict__ y, int64_t n, uint16_t *dz) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float g = dy[i];
        if (y && !(bf2f(y[i]) > 0.f)) g = 0.f;
        dz[i] = f2bf(g);
    }
}

// Mean squared error over n elements: loss partials and grad = 2 (o - t) / n.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ o, const float *__restrict__ t, int64_t n,
                                             float *__restrict__ grad, float *__restrict__ partial) {
    __shared__ float red[256];
    float acc = 0.f;
    const float inv = 2.f / (float)n;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float d = o[i] - t[i];
        acc += d * d;
        grad[i] = inv * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam on FP32 master weights, refreshing the BF16 copy used by the kernels.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, uint16_t *__restrict__ wb, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, float wd) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i] + wd * w[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        const float v = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
        w[i] = v;
        if (wb) wb[i] = f2bf(v);
    }
}

// fp32 -> bf16 copy.
__global__ void k_to_bf16(const float *__restrict__ s, int64_t n, uint16_t *__restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = f2bf(s[i]);
}

// wf[ci][kh][kw][co] = w[co][2-kh][2-kw][ci]: the flipped, transposed kernel
// whose forward convolution is the input gradient of w.
__global__ void k_flip_weights(const uint16_t *__restrict__ w, int Cout, int Cin, uint16_t *__restrict__ wf) {
    const int64_t n = (int64_t)Cout * 9 * Cin;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ci = (int)(i % Cin);
        const int tap = (int)((i / Cin) % 9);
        const int co = (int)(i / (Cin * 9));
        const int kh = tap / 3, kw = tap % 3;
        wf[((int64_t)ci * 9 + (2 - kh) * 3 + (2 - kw)) * Cout + co] = w[i];
    }
}

// Warm-start input: costs [L][H][W] (reference layout) min-pooled over f x f
// blocks, scaled, written NHWC bf16 [h][w][L] for instance slot `n`.  One
// thread per pooled cell with X fastest, so a wave reads 64 adjacent f-wide
// row segments of one label plane (coalesced); the bf16 writes are strided by
// L but the output is f*f*2 times smaller than the input.
template <typename T>
__global__ void k_minpool(const T *__restrict__ d, int W, int H, int L, int f, float scale, int w4, int h4,
                          uint16_t *__restrict__ out, bool vec4) {
    const int64_t n = (int64_t)w4 * h4 * L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int X = (int)(i % w4);
        const int64_t r = i / w4;
        const int Y = (int)(r % h4);
        const int l = (int)(r / h4);
        const int x1 = min(X * f + f, W), y1 = min(Y * f + f, H);
        const T *plane = d + (int64_t)l * H * W;
        float mn = 3.4e38f;
        if (f == 4 && x1 - X * f == 4 && vec4) {
            // whole 4-wide row segments, 16-byte aligned: one load per row,
            // the same left-to-right comparisons
            typedef T t4 __attribute__((ext_vector_type(4)));
            for (int yy = Y * f; yy < y1; ++yy) {
                const t4 v4 = *reinterpret_cast<const t4 *>(plane + (int64_t)yy * W + X * f);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float v = (float)v4[k];
                    mn = v < mn ? v : mn;
                }
            }
        } else {
            for (int yy = Y * f; yy < y1; ++yy)
                for (int xx = X * f; xx < x1; ++xx) {
                    const float v = (float)plane[(int64_t)yy * W + xx];
                    mn = v < mn ? v : mn;
                }
        }
        out[((int64_t)Y * w4 + X) * L + l] = f2bf(mn * scale);
    }
}

// The same pooling for 4x4 blocks of whole 16-byte rows and L % 8 == 0: one
// thread per (cell, group of 8 labels) with X fastest, 32 row loads in
// flight, the same left-to-right comparisons per label, and one 16-byte
// store of the 8 labels' bf16 values (the per-label kernel's 2-byte stores
// strided by L each touched a line of their own).
template <typename T>
__global__ void k_minpool8(const T *__restrict__ d, int W, int H, int L, float scale, int w4, int h4,
                           uint16_t *__restrict__ out) {
    typedef T t4 __attribute__((ext_vector_type(4)));
    const int64_t n = (int64_t)w4 * h4 * (L / 8);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int X = (int)(i % w4);
        const int64_t r = i / w4;
        const int Y = (int)(r % h4);
        const int lg = (int)(r / h4);
        // All 32 row reads (8 labels x 4 rows) are issued before the first
        // minimum: the rows are unrolled with a guard at the bottom edge
        // instead of a loop bound that would serialise them.
        t4 v4[8][4];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const T *plane = d + (int64_t)(8 * lg + j) * H * W + X * 4;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (Y * 4 + r < H) v4[j][r] = *reinterpret_cast<const t4 *>(plane + (int64_t)(Y * 4 + r) * W);
        }
        uint32_t packed[4];
#pragma unroll
        for (int k = 0; k < 8; k += 2) {
            float mn[2] = {3.4e38f, 3.4e38f};
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (Y * 4 + r >= H) continue;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float v = (float)v4[k + j][r][e];
                        mn[j] = v < mn[j] ? v : mn[j];
                    }
                }
            packed[k / 2] = (uint32_t)f2bf(mn[0] * scale) | ((uint32_t)f2bf(mn[1] * scale) << 16);
        }
        *reinterpret_cast<uint4 *>(out + ((int64_t)Y * w4 + X) * L + 8 * lg) =
            uint4{packed[0], packed[1], packed[2], packed[3]};
    }
}

// Training target: level-0 outboxes (oracle layout [l][d][y][x]) averaged over
// f x f blocks, scaled, NHWC f32 [h][w][4L] with channel d*L + l.
__global__ void k_block_mean_msgs(const float *__restrict__ m, int W, int H, int L, int f, float scale, int w4,
                                  int h4, float *__restrict__ out) {
    const int64_t n = (int64_t)w4 * h4 * 4 * L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % (4 * L));
        const int X = (int)((i / (4 * L)) % w4);
        const int Y = (int)(i / ((int64_t)4 * L * w4));
        const int d = ch / L, l = ch % L;
        float acc = 0.f;
        int cnt = 0;
        for (int yy = Y * f; yy < Y * f + f && yy < H; ++yy)
            for (int xx = X * f; xx < X * f + f && xx < W; ++xx) {
                acc += m[(((int64_t)l * 4 + d) * H + yy) * W + xx];
                ++cnt;
            }
        out[i] = cnt ? acc / cnt * scale : 0.f;
    }
}

// Seed the outboxes of one plan level from a prediction at 1/2^shift of its
// resolution: interior node (x, y) gets pred[(y >> shift)][(x >> shift)][d*L + l]
// / scale; border outboxes stay 0.
struct InjectArgs {
    void *msgs[2][4];
    ginf::LevelGeom g;
    int C, G, interleave, L, shift, pw, ph;
    float inv_scale;
    int pred16;  // prediction rows 16-byte aligned (L % 4 == 0, aligned base)
};

// One thread per slot of the outbox array of colour blockIdx.y & 1 and
// direction blockIdx.y >> 1, looping over the register planes: each store
// writes that slot's G labels contiguously (a 16-byte store for the quad
// layout), so the writes of a wave (64 slots of one slot tile) are one
// contiguous run per plane, and the successive planes read successive 16 bytes
// of the same prediction row (one cache line serves eight planes).  Border
// nodes get 0; padding slots and padding labels are left as they are.
__global__ void k_inject(const float *__restrict__ pred, InjectArgs a) {
    const int c = blockIdx.y & 1, d = blockIdx.y >> 1;
    float *__restrict__ out = (float *)a.msgs[c][d];
    const int G = a.G, C = a.C, L = a.L;
    typedef float f4v __attribute__((ext_vector_type(4)));
    for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < a.g.nslots; slot += gridDim.x * blockDim.x) {
        const int y = slot / a.g.whp;
        if (y >= a.g.h) continue;
        const int x = 2 * (slot - y * a.g.whp) + ((c ^ y) & 1);
        if (x >= a.g.w) continue;
        const bool interior = x >= 1 && y >= 1 && x <= a.g.w - 2 && y <= a.g.h - 2;
        const int X = min(x >> a.shift, a.pw - 1), Y = min(y >> a.shift, a.ph - 1);
        const float *p = pred + ((int64_t)Y * a.pw + X) * 4 * L + d * L;
        for (int ii = 0; ii < C; ++ii) {
            float *o = out + ginf::elem_index(slot, ii, 0, C, G);
            if (a.interleave && G == 4 && ii * 4 + 3 < L) {
                f4v v = {0.f, 0.f, 0.f, 0.f};
                if (interior) {
                    if (a.pred16) {  // L % 4 == 0 and an aligned prediction: one 16-byte read
                        v = *reinterpret_cast<const f4v *>(p + ii * 4);
                    } else {
                        v[0] = p[ii * 4 + 0];
                        v[1] = p[ii * 4 + 1];
                        v[2] = p[ii * 4 + 2];
                        v[3] = p[ii * 4 + 3];
                    }
                    v[0] *= a.inv_scale;
                    v[1] *= a.inv_scale;
                    v[2] *= a.inv_scale;
                    v[3] *= a.inv_scale;
                }
                __builtin_nontemporal_store(v, reinterpret_cast<f4v *>(o));  // not read again in this launch
            } else {
                for (int q = 0; q < G; ++q) {
                    const int l = a.interleave ? ii * G + q : q * C + ii;
                    if (l < L) o[q] = interior ? p[l] * a.inv_scale : 0.f;
                }
            }
        }
    }
}

// Block mean of a plan level's outboxes, read from the state layout.
__global__ void k_block_mean_state(InjectArgs a, int f, float scale, float *__restrict__ out) {
    const int w4 = (a.g.w + f - 1) / f, h4 = (a.g.h + f - 1) / f;
    const int64_t n = (int64_t)w4 * h4 * 4 * a.L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % (4 * a.L));
        const int X = (int)((i / (4 * a.L)) % w4);
        const int Y = (int)(i / ((int64_t)4 * a.L * w4));
        const int d = ch / a.L, l = ch % a.L;
        const int q = a.interleave ? l % a.G : l / a.C;
        const int ii = a.interleave ? l / a.G : l - q * a.C;
        float acc = 0.f;
        int cnt = 0;
        for (int y = Y * f; y < Y * f + f && y < a.g.h; ++y)
            for (int x = X * f; x < X * f + f && x < a.g.w; ++x) {
                const int c = (x + y) & 1;
                const int64_t slot = (int64_t)y * a.g.whp + (x >> 1);
                acc += ((const float *)a.msgs[c][d])[ginf::elem_index(slot, ii, q, a.C, a.G)];
                ++cnt;
            }
        out[i] = cnt ? acc / cnt * scale : 0.f;
    }
}

// Stereo data costs from an image pair, the cost rule of ginf_gen_stereo:
// min(|left(x,y) - right(x-l,y)|, trunc), trunc where x - l < 0.
__global__ void k_stereo_costs(const uint8_t *__restrict__ left, const uint8_t *__restrict__ right, int W, int H,
                               int L, int trunc, float *__restrict__ out) {
    const int64_t n = (int64_t)W * H * L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % W);
        const int y = (int)((i / W) % H);
        const int l = (int)(i / ((int64_t)W * H));
        int c = trunc;
        if (x - l >= 0) {
            c = (int)left[(int64_t)y * W + x] - (int)right[(int64_t)y * W + x - l];
            c = c < 0 ? -c : c;
            c = c < trunc ? c : trunc;
        }
        out[i] = (float)c;
    }
}

inline unsigned nblk(int64_t n, int per = 256) {
    int64_t b = (n + per - 1) / per;
    return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

}  // namespace

extern "C" {

int ginf_conv3x3_fwd(const uint16_t *x, const uint16_t *w, const float *bias, int N, int H, int W, int Cin, int Cout,
                     int relu, uint16_t *y_bf16, float *y_f32, void *stream) {
    if (!x || !w || (!y_bf16 && !y_f32) || N < 1 || H < 1 || W < 1 || Cin % 32 || Cout % 32 || Cout < 32)
        return GINF_ERR_INVALID;
    const int64_t P = (int64_t)N * H * W;
    const dim3 grid((unsigned)((P + 63) / 64), (unsigned)(Cout / 32)), block(256);
    hipStream_t st = (hipStream_t)stream;
#define GINF_CONV_PF(CS, TAPS) \
    hipLaunchKernelGGL((k_conv3x3_pf<CS, TAPS>), grid, block, 0, st, x, w, bias, N, H, W, Cout, relu, y_bf16, y_f32)
    switch (Cin / 32) {
        case 1: GINF_CONV_PF(1, 9); break;
        case 2: GINF_CONV_PF(2, 3); break;
        case 3: GINF_CONV_PF(3, 3); break;
        case 4: GINF_CONV_PF(4, 3); break;
        case 5: GINF_CONV_PF(5, 1); break;
        case 6: GINF_CONV_PF(6, 1); break;
        case 7: GINF_CONV_PF(7, 1); break;
        case 8: GINF_CONV_PF(8, 1); break;
        default:
            hipLaunchKernelGGL(k_conv3x3, grid, block, 0, st, x, w, bias, N, H, W, Cin, Cout, relu, y_bf16, y_f32);
    }
#undef GINF_CONV_PF
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

int ginf_convnet_fwd(int layers, const int *chans, const uint16_t *const *w, const float *const *b, const uint16_t *x,
                     int N, int H, int W, uint16_t *act0, uint16_t *act1, float *out, void *stream) {
    if (layers < 1 || !chans || !w || !b || !x || !out) return GINF_ERR_INVALID;
    // The shapes the one-launch kernel is built for run there (same results).
    const hipError_t fe = ginf::launch_convnet_fused(layers, chans, w, b, x, N, H, W, out, (hipStream_t)stream);
    if (fe == hipSuccess) return GINF_OK;
    if (fe != hipErrorNotSupported) return GINF_ERR_HIP;
    if (layers > 1 && (!act0 || !act1)) return GINF_ERR_INVALID;
    for (int k = 0; k < layers; ++k) {
        const uint16_t *in = k == 0 ? x : (k % 2 ? act0 : act1);
        const bool last = k == layers - 1;
        uint16_t *yb = last ? nullptr : (k % 2 ? act1 : act0);
        const int rc = ginf_conv3x3_fwd(in, w[k], b[k], N, H, W, chans[k], chans[k + 1], last ? 0 : 1, yb,
                                        last ? out : nullptr, stream);
        if (rc != GINF_OK) return rc;
    }
    return GINF_OK;
}

int ginf_conv3x3_bwd_weight(const uint16_t *x, const uint16_t *dz, int N, int H, int W, int Cin, int Cout, float *dw,
                            float *db, void *stream) {
    if (!x || !dz || !dw || Cin % 16 || Cout % 16) return GINF_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    const unsigned blocks = (unsigned)((Cout / 16) * (Cin / 16) * 9);
    const int64_t P = (int64_t)N * H * W, nout = (int64_t)Cout * 9 * Cin;
    // about 32 chunks of whole 128-pixel rounds (one round = the block's four waves x 32 pixels)
    int64_t chunk = ((P + 31) / 32 + 127) / 128 * 128;
    if (chunk < 128) chunk = 128;
    const int nchunk = (int)((P + chunk - 1) / chunk);
    const int cw = Cout % 64 == 0 ? 64 : (Cout % 32 == 0 ? 32 : 16);
    // the bias gradient is a much smaller sum per block: more, shorter chunks
    const int64_t bchunk = std::max<int64_t>(64, (P + 255) / 256);
    const int nbchunk = (int)((P + bchunk - 1) / bchunk);
    float *part = nullptr;
    // weight partials for up to max(nchunk, kDwBlocks) blocks, then the bias partials
    const int64_t wparts = std::max<int64_t>(nchunk, kDwBlocks);
    const size_t part_n = (size_t)wparts * (size_t)nout + (size_t)nbchunk * (size_t)Cout;
    if (hipMallocAsync((void **)&part, part_n * sizeof(float), st) != hipSuccess) return GINF_ERR_NOMEM;
    const int cit = Cin / 16;
    // co tiles per block: the most (of 4, 2, 1) dividing Cout / 16 whose
    // accumulators stay within 80 registers per lane
    int ct = 1;
    for (int c : {4, 2})
        if ((Cout / 16) % c == 0 && c * ((9 * cit + 3) / 4) <= 20) { ct = c; break; }
    if (cit == 1 || cit == 2 || cit == 3 || cit == 4 || cit == 6 || cit == 8) {
        // one pixel block per CU and co group, every block walking its share
        // of the 4x32 patches
        const int64_t npatch = (int64_t)N * ((H + kDwPR - 1) / kDwPR) * ((W + kDwPC - 1) / kDwPC);
        const int ngroups = Cout / (16 * ct);
        const int npb = (int)std::max<int64_t>(1, std::min<int64_t>(npatch, std::max(8, kDwBlocks / ngroups)));
        const dim3 grid((unsigned)npb, (unsigned)ngroups);
        float *pw = npb > 1 ? part : dw;
#define GINF_DWT(CI, CT) \
    if (cit == CI && ct == CT) hipLaunchKernelGGL((k_conv3x3_dw_tr<CI, CT>), grid, dim3(256), 0, st, x, dz, N, H, W, Cout, pw)
        GINF_DWT(1, 4); GINF_DWT(1, 2); GINF_DWT(1, 1); GINF_DWT(2, 4); GINF_DWT(2, 2); GINF_DWT(2, 1);
        GINF_DWT(3, 2); GINF_DWT(3, 1); GINF_DWT(4, 2); GINF_DWT(4, 1); GINF_DWT(6, 1); GINF_DWT(8, 1);
#undef GINF_DWT
        if (npb > 1)
            hipLaunchKernelGGL(k_chunk_reduce, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, part, npb, nout,
                               dw);
    } else {
        float *pw = nchunk > 1 ? part : dw;
        hipLaunchKernelGGL(k_conv3x3_dw, dim3(blocks, (unsigned)nchunk), dim3(256), 0, st, x, dz, N, H, W, Cin, Cout,
                           chunk, pw);
        if (nchunk > 1)
            hipLaunchKernelGGL(k_chunk_reduce, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, part, nchunk,
                               nout, dw);
    }
    if (db) {
        float *pb = part + (size_t)wparts * nout;
        hipLaunchKernelGGL(k_bias_grad_part, dim3((unsigned)(Cout / cw), (unsigned)nbchunk), dim3(256), 0, st, dz, P,
                           Cout, cw, bchunk, pb);
        hipLaunchKernelGGL(k_chunk_reduce, dim3((unsigned)((Cout + 255) / 256)), dim3(256), 0, st, pb, nbchunk,
                           (int64_t)Cout, db);
    }
    if (hipFreeAsync(part, st) != hipSuccess) return GINF_ERR_HIP;
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

int ginf_conv3x3_flip_weights(const uint16_t *w, int Cout, int Cin, uint16_t *wf, void *stream) {
    if (!w || !wf) return GINF_ERR_INVALID;
    hipLaunchKernelGGL(k_flip_weights, dim3(nblk((int64_t)Cout * 9 * Cin)), dim3(256), 0, (hipStream_t)stream, w,
                       Cout, Cin, wf);
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

int ginf_relu_bwd(const float *dy, const uint16_t *y, int64_t n, uint16_t *dz, void *stream) {
    if (!dy || !dz) return GINF_ERR_INVALID;
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, dy, y, n, dz);
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

int ginf_mse(const float *o, const float *t, int64_t n, float *grad, float *partials /* >= 1024 */, int *nparts,
             void *stream) {
    if (!o || !t || !grad || !partials || !nparts) return GINF_ERR_INVALID;
    const unsigned b = nblk(n) > 1024 ? 1024 : nblk(n);
    hipLaunchKernelGGL(k_mse, dim3(b), dim3(256), 0, (hipStream_t)stream, o, t, n, grad, partials);
    *nparts = (int)b;
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

int ginf_adam(float *w, const float *g, float *m1, float *m2, uint16_t *wb, int64_t n, float lr, float b1, float b2,
              float eps, int step, float wd, void *stream) {
    if (!w || !g || !m1 || !m2 || step < 1) return GINF_ERR_INVALID;
    const float c1 = 1.f - __builtin_powf(b1, (float)step), c2 = 1.f - __builtin_powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, w, g, m1, m2, wb, n, lr, b1, b2, eps,
                       c1, c2, wd);
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

int ginf_to_bf16(const float *s, int64_t n, uint16_t *d, void *stream) {
    if (!s || !d) return GINF_ERR_INVALID;
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, s, n, d);
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

int ginf_minpool_costs(const void *costs, int dtype, int W, int H, int L, int f, float scale, uint16_t *out,
                       void *stream) {
    if (!costs || !out || f < 1) return GINF_ERR_INVALID;
    const int w4 = (W + f - 1) / f, h4 = (H + f - 1) / f;
    const int64_t n = (int64_t)w4 * h4 * L;
    // 16-byte row segments: every row and plane starts 16-byte aligned
    const bool vec4 = f == 4 && W % 4 == 0 && ((uintptr_t)costs & 15) == 0;
    if (vec4 && L % 8 == 0 && ((uintptr_t)out & 15) == 0) {
        const int64_t n8 = (int64_t)w4 * h4 * (L / 8);
        if (dtype == GINF_F32)
            hipLaunchKernelGGL(k_minpool8<float>, dim3(nblk(n8)), dim3(256), 0, (hipStream_t)stream,
                               (const float *)costs, W, H, L, scale, w4, h4, out);
        else
            hipLaunchKernelGGL(k_minpool8<int>, dim3(nblk(n8)), dim3(256), 0, (hipStream_t)stream, (const int *)costs,
                               W, H, L, scale, w4, h4, out);
        return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
    }
    if (dtype == GINF_F32)
        hipLaunchKernelGGL(k_minpool<float>, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, (const float *)costs, W,
                           H, L, f, scale, w4, h4, out, vec4);
    else
        hipLaunchKernelGGL(k_minpool<int>, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, (const int *)costs, W, H,
                           L, f, scale, w4, h4, out, vec4);
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

int ginf_stereo_costs(const uint8_t *left, const uint8_t *right, int W, int H, int L, int trunc, float *out,
                      void *stream) {
    if (!left || !right || !out || W < 1 || H < 1 || L < 1) return GINF_ERR_INVALID;
    hipLaunchKernelGGL(k_stereo_costs, dim3(nblk((int64_t)W * H * L)), dim3(256), 0, (hipStream_t)stream, left, right,
                       W, H, L, trunc, out);
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

int ginf_block_mean_msgs(const float *msgs, int W, int H, int L, int f, float scale, float *out, void *stream) {
    if (!msgs || !out || f < 1) return GINF_ERR_INVALID;
    const int w4 = (W + f - 1) / f, h4 = (H + f - 1) / f;
    const int64_t n = (int64_t)w4 * h4 * 4 * L;
    hipLaunchKernelGGL(k_block_mean_msgs, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, msgs, W, H, L, f, scale,
                       w4, h4, out);
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

}  // extern "C"

// Used by bp_plan.cpp (plan internals stay there).
namespace ginf {
hipError_t launch_inject(const float *pred, void *const msgs[2][4], const LevelGeom &g, int C, int G, bool interleave,
                         int L, int shift, int pw, int ph, float inv_scale, hipStream_t st) {
    InjectArgs a{};
    for (int c = 0; c < 2; ++c)
        for (int d = 0; d < 4; ++d) a.msgs[c][d] = msgs[c][d];
    a.g = g;
    a.C = C;
    a.G = G;
    a.interleave = interleave ? 1 : 0;
    a.L = L;
    a.shift = shift;
    a.pw = pw;
    a.ph = ph;
    a.inv_scale = inv_scale;
    a.pred16 = (L % 4 == 0 && ((uintptr_t)pred & 15) == 0) ? 1 : 0;
    hipLaunchKernelGGL(k_inject, dim3(nblk(g.nslots), 8), dim3(256), 0, st, pred, a);
    return hipGetLastError();
}

hipError_t launch_block_mean_state(void *const msgs[2][4], const LevelGeom &g, int C, int G, bool interleave, int L,
                                   int f, float scale, float *out, hipStream_t st) {
    InjectArgs a{};
    for (int c = 0; c < 2; ++c)
        for (int d = 0; d < 4; ++d) a.msgs[c][d] = msgs[c][d];
    a.g = g;
    a.C = C;
    a.G = G;
    a.interleave = interleave ? 1 : 0;
    a.L = L;
    const int64_t n = (int64_t)((g.w + f - 1) / f) * ((g.h + f - 1) / f) * 4 * L;
    hipLaunchKernelGGL(k_block_mean_state, dim3(nblk(n)), dim3(256), 0, st, a, f, scale, out);
    return hipGetLastError();
}
}  // namespace ginf
