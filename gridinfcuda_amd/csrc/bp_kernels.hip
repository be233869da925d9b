// HIP kernels (gfx950) around the half-sweeps: data-cost relayout and
// pyramid, explicit outbox upsample, fills, labels and the FP64 energy, plus
// the half-sweep dispatch (kernels in sweep_*.hip).  Layout: see bp_layout.h.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bp_kernels.h"
#include "conv_common.h"

namespace ginf {
namespace {

// Reference-layout costs [l][y][x] -> level-0 device layout, both colours.
// Labels >= L (padding up to C*G) get cost 0; the sweeps mask them.
template <typename T>
__global__ __launch_bounds__(256) void k_relayout(const T *__restrict__ src, int W, int H, int L, StateLayout ly,
                                                  LevelGeom g, T *dst0, T *dst1) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    T *dst = ((x + y) & 1) ? dst1 : dst0;
    const int64_t slot = (int64_t)y * g.whp + (x >> 1);
    const int LP = ly.C * ly.G;
    for (int l = 0; l < LP; ++l) {
        const int qq = ly.interleave ? l % ly.G : l / ly.C;
        const int i = ly.interleave ? l / ly.G : l - qq * ly.C;
        const T v = (l < L) ? src[((int64_t)l * H + y) * W + x] : T(0);
        dst[elem_index(slot, i, qq, ly.C, ly.G)] = v;
    }
}

// Relayout fused with the first pyramid level.  A block takes fine rows
// 2Y and 2Y+1 and TW consecutive columns of the reference-layout costs
// [l][y][x], stages them in LDS (coalesced reads along x), then writes the
// level-0 state (both colours; a plane's slots are contiguous within a slot tile) and,
// when `coarse` is set, the level-1 costs of coarse row Y: the children
// (2X,2Y), (2X+1,2Y), (2X,2Y+1), (2X+1,2Y+1) added in that order from 0
// (src/bp.cc:100-117).  Padding labels (L <= l < C*G) get cost 0.
// Widest label count (C*G) whose relayout tile is 64 columns wide (else 32):
// the tile holds 2 rows x C*G labels x (TW+1) values in LDS.
#ifndef GINF_RELAYOUT_TW64_MAX_LP
#define GINF_RELAYOUT_TW64_MAX_LP 64
#endif

// Storage of a data cost: T itself, or an unsigned integer when S is
// uint16_t or uint8_t.  to_store records in *bad when a value has no exact
// form in S (not an integer in [0, 65535] or [0, 255], or -0), and the plan
// then redoes the level data in the next wider form.
template <typename T, int MAXV>
__device__ __forceinline__ bool fits_uint(T v) {
    if constexpr (std::is_same<T, float>::value)
        return v >= T(0) && v <= T(MAXV) && v == __builtin_truncf(v) && !__builtin_signbit(v);
    else
        return v >= T(0) && v <= T(MAXV);
}

template <typename S, typename T>
__device__ __forceinline__ S to_store(T v, bool &bad) {
    if constexpr (sizeof(S) < sizeof(T)) {
        constexpr int MAXV = sizeof(S) == 1 ? 255 : 65535;
        const bool ok = fits_uint<T, MAXV>(v);
        bad |= !ok;
        return (S)(uint32_t)(ok ? v : T(0));
    } else {
        return v;
    }
}

template <typename S, typename T>
__device__ __forceinline__ T from_store(S v) {
    if constexpr (sizeof(S) < sizeof(T)) return T((uint32_t)v);
    else return v;
}

__device__ __forceinline__ void report_bad(bool bad, int *flag) {
    if (bad && flag) atomicOr(flag, 1);
}

template <typename T, typename S, int TW>
__global__ __launch_bounds__(256) void k_relayout_pyr(const T *__restrict__ src, int W, int H, int L, StateLayout ly,
                                                      LevelGeom g0, S *d00, S *d01, int coarse, LevelGeom g1, S *d10,
                                                      S *d11, int *flag, const int *only_if) {
    if (only_if && *only_if == 0) return;  // conditional redo in T: nothing to redo
    bool bad = false;
    extern __shared__ unsigned char smem_raw[];
    constexpr int TP = TW + 1;  // padded row: label-strided LDS reads hit distinct banks
    T *tile = reinterpret_cast<T *>(smem_raw);  // [2][LP][TP]
    const int G = ly.G, C = ly.C;
    const int gs = __builtin_ctz(G);  // G is a power of two
    const int LP = C * G;
    const int x0 = blockIdx.x * TW;
    const int Y = blockIdx.y;
    const int tid = threadIdx.x;
    // stage: 2 rows x LP labels x TW columns, columns fastest; eight loads
    // in flight per thread before their LDS writes
    {
        constexpr int STEP = 256 / TW;
        const int xl = tid % TW;
        const int x = x0 + xl;
        for (int base = tid / TW; base < 2 * LP; base += 8 * STEP) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int rl = base + u * STEP;
                const int rr = rl >= LP ? 1 : 0;
                const int l = rl - rr * LP;
                const int y = 2 * Y + rr;
                v[u] = (rl < 2 * LP && l < L && x < W && y < H) ? src[((int64_t)l * H + y) * W + x] : T(0);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int rl = base + u * STEP;
                if (rl < 2 * LP) tile[rl * TP + xl] = v[u];
            }
        }
    }
    __syncthreads();
    // level 0: per (row, colour, plane i): HS slots x G lanes, contiguous within a slot tile
    constexpr int HS = TW / 2;
    {
        const int per = HS << gs;          // elements per (row, colour, plane)
        const int ni = 256 / per > 0 ? 256 / per : 1;
        const int sub = tid / per, r = tid % per;
        const int q = r & (G - 1), k = r >> gs;
        if (sub < ni) {
            for (int rc = 0; rc < 4; ++rc) {
                const int rr = rc >> 1, c = rc & 1;
                const int y = 2 * Y + rr;
                const int xl = 2 * k + ((y + c) & 1);
                const int x = x0 + xl;
                if (y >= H || x >= W) continue;
                const int64_t slot = (int64_t)y * g0.whp + (x >> 1);
                S *dst = c ? d01 : d00;
                for (int i = sub; i < C; i += ni) {
                    const int l = ly.interleave ? (i << gs) + q : q * C + i;
                    dst[elem_index(slot, i, q, C, G)] = to_store<S>(tile[(rr * LP + l) * TP + xl], bad);
                }
            }
        }
    }
    if (!coarse) {
        report_bad(bad, flag);
        return;
    }
    // level 1: coarse row Y, QS slots per colour
    constexpr int QS = TW / 4;
    {
        const int per = QS << gs;
        const int ni = 256 / per > 0 ? 256 / per : 1;
        const int sub = tid / per, r = tid % per;
        const int q = r & (G - 1), k = r >> gs;
        if (sub >= ni) {
            report_bad(bad, flag);
            return;
        }
        for (int c = 0; c < 2; ++c) {
            const int Xl = 2 * k + ((Y + c) & 1);
            const int X = x0 / 2 + Xl;
            if (X >= g1.w) continue;
            const bool ex = 2 * X + 1 < W, ey = 2 * Y + 1 < H;
            const int64_t slot = (int64_t)Y * g1.whp + (X >> 1);
            S *dst = c ? d11 : d10;
            for (int i = sub; i < C; i += ni) {
                const int l = ly.interleave ? (i << gs) + q : q * C + i;
                const T *t0 = tile + l * TP;
                const T *t1 = tile + (LP + l) * TP;
                T acc = T(0);
                acc = acc + t0[2 * Xl];
                if (ex) acc = acc + t0[2 * Xl + 1];
                if (ey) acc = acc + t1[2 * Xl];
                if (ex && ey) acc = acc + t1[2 * Xl + 1];
                dst[elem_index(slot, i, q, C, G)] = to_store<S>(acc, bad);
            }
        }
    }
    report_bad(bad, flag);
}

// Data-cost pyramid level (src/bp.cc:100-117): parent (X,Y) sums its children
// (2X,2Y), (2X+1,2Y), (2X,2Y+1), (2X+1,2Y+1) in that order, starting from 0.
// Children (2X,2Y) and (2X+1,2Y+1) have colour 0, the other two colour 1, and
// all four sit in slot column X of their rows.
// One thread per (plane i, coarse node, lane q): four loads and one store,
// so the small coarse levels are not a chain of C dependent plane steps.
template <typename T, typename S>
__global__ __launch_bounds__(256) void k_pyramid(const S *f0, const S *f1, LevelGeom fg, S *c0, S *c1,
                                                 LevelGeom cg, int C, int G, int *flag, const int *only_if) {
    if (only_if && *only_if == 0) return;  // conditional redo in T: nothing to redo
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per_plane = (int64_t)cg.w * cg.h * G;
    const int i = (int)(gid / per_plane);
    if (i >= C) return;
    const int r = (int)(gid - (int64_t)i * per_plane);
    const int node = r / G, q = r - node * G;
    const int X = node % cg.w, Y = node / cg.w;
    const bool ex = 2 * X + 1 < fg.w, ey = 2 * Y + 1 < fg.h;
    const int64_t s00 = (int64_t)(2 * Y) * fg.whp + X;
    const int64_t s01 = (int64_t)(2 * Y + 1) * fg.whp + X;
    S *dst = ((X + Y) & 1) ? c1 : c0;
    const int64_t cs = (int64_t)Y * cg.whp + (X >> 1);
    bool bad = false;
    const int64_t e00 = elem_index(s00, i, q, C, G), e01 = elem_index(s01, i, q, C, G);
    T acc = T(0);
    acc = acc + from_store<S, T>(f0[e00]);
    if (ex) acc = acc + from_store<S, T>(f1[e00]);
    if (ey) acc = acc + from_store<S, T>(f1[e01]);
    if (ex && ey) acc = acc + from_store<S, T>(f0[e01]);
    dst[elem_index(cs, i, q, C, G)] = to_store<S>(acc, bad);
    report_bad(bad, flag);
}

// Explicit outbox upsample (src/bp.cc:129-139), every fine node including borders.
template <typename T>
__global__ __launch_bounds__(256) void k_upsample(UpsampleArgs u) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int G = u.G;
    const int node = gid / G, q = gid - (gid / G) * G;
    const int x = node % u.fg.w, y = node / u.fg.w;
    if (y >= u.fg.h) return;
    const int px = x >> 1, py = y >> 1;
    const int fc = (x + y) & 1, pc = (px + py) & 1;
    const int64_t fs = (int64_t)y * u.fg.whp + (x >> 1);
    const int64_t ps = (int64_t)py * u.cg.whp + (px >> 1);
    for (int d = 0; d < 4; ++d) {
        const T *src = (const T *)u.coarse[pc][d];
        T *dst = (T *)u.fine[fc][d];
        for (int i = 0; i < u.C; ++i)
            dst[elem_index(fs, i, q, u.C, G)] = src[elem_index(ps, i, q, u.C, G)];
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Interior nodes of colour `color` get label 0 (their belief stayed zero).
__global__ __launch_bounds__(256) void k_zero_labels(int32_t *labels, int W, int H, int color) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x < 1 || x > W - 2 || y < 1 || y > H - 2) return;
    if (((x + y) & 1) == color) labels[(int64_t)y * W + x] = 0;
}

// FP64 energy: per-block partial sums over interior nodes in row-major order,
// then one block adds the partials in index order (deterministic).
template <typename T>
__global__ __launch_bounds__(256) void k_energy_partial(const T *data, const T *table, const int32_t *f, int W,
                                                        int H, int L, double *partials) {
    __shared__ double red[256];
    // Grids without interior nodes (W < 3 or H < 3) have energy 0
    // (src/grid.cc:102-125 sums over 1 <= x <= W-2, 1 <= y <= H-2); the clamp
    // keeps (W-2)*(H-2) from turning positive for W = H = 1.
    const int64_t nin = (W < 3 || H < 3) ? 0 : (int64_t)(W - 2) * (H - 2);
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nin; i += (int64_t)gridDim.x * blockDim.x) {
        const int y = 1 + (int)(i / (W - 2));
        const int x = 1 + (int)(i - (int64_t)(y - 1) * (W - 2));
        const int fl = f[(int64_t)y * W + x];
        const int fn = f[(int64_t)(y - 1) * W + x];
        const int fe = f[(int64_t)y * W + x + 1];
        if ((unsigned)fl >= (unsigned)L || (unsigned)fn >= (unsigned)L || (unsigned)fe >= (unsigned)L) {
            acc = __builtin_nan("");  // label out of range: reported as an error by the host
            continue;
        }
        acc += (double)data[((int64_t)fl * H + y) * W + x];
        acc += (double)table[fl + L * fn];
        acc += (double)table[fl + L * fe];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_energy_final(const double *partials, int n, double *out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += partials[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}


// Relayout for the label-quad layout (l = 4i + q), without LDS: block
// (column-pair run, row pair Y, quad i); thread k takes columns 2k, 2k+1 of
// rows 2Y, 2Y+1 for labels 4i..4i+3 (eight 8-byte loads from rows that a wave
// reads as 512-byte runs), writes the four level-0 nodes as one quad each
// (slot k of row y in the node's colour) and, fused, the level-1 node (k, Y)
// as the child sum in the reference order (2X,2Y), (2X+1,2Y), (2X,2Y+1),
// (2X+1,2Y+1).  Labels past L are stored as 0, as the tiled kernel does.
// Planes i = blockIdx.z, blockIdx.z + gridDim.z, ... < C: the plan launches
// one z-block per plane, and the conditional redo one z-block in all, so that
// the redo's blocks are few when they have nothing to do.
// VEC (even W, 8-byte aligned source): the two columns of a row are one
// 8-byte load.
template <typename T, typename S, bool VEC>
__device__ __forceinline__ void relayout_quad_pair(const T *__restrict__ src, int W, int H, int L, int C,
                                                   const LevelGeom &g0, S *d00, S *d01, int coarse, const LevelGeom &g1,
                                                   S *d10, S *d11, int k, int Y, int i, bool &bad, T (&v)[4][2][2]) {
    const int x0 = 2 * k, y0 = 2 * Y;
    const bool ex = x0 + 1 < W, ey = y0 + 1 < H;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int l = 4 * i + q;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            v[q][r][0] = T(0);
            v[q][r][1] = T(0);
            if (l < L && (r == 0 || ey)) {
                const T *row = src + ((int64_t)l * H + y0 + r) * W;
                if constexpr (VEC) {
                    typedef T t2 __attribute__((ext_vector_type(2)));
                    const t2 p = *reinterpret_cast<const t2 *>(row + x0);  // W even: x0 + 1 < W
                    v[q][r][0] = p[0];
                    v[q][r][1] = p[1];
                } else {
                    v[q][r][0] = row[x0];
                    if (ex) v[q][r][1] = row[x0 + 1];
                }
            }
        }
    }
    typedef S s4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        if (r == 1 && !ey) break;
        const int y = y0 + r;
        const int64_t slot = (int64_t)y * g0.whp + k;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (c == 1 && !ex) break;
            const int x = x0 + c;
            S *dst = ((x + y) & 1) ? d01 : d00;
            s4 o;
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = to_store<S>(v[q][r][c], bad);
            *(s4 *)(dst + elem_index(slot, i, 0, C, 4)) = o;
        }
    }
    if (coarse && k < g1.w) {
        const int64_t slot = (int64_t)Y * g1.whp + (k >> 1);
        S *dst = ((k + Y) & 1) ? d11 : d10;
        s4 o;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T acc = T(0);
            acc = acc + v[q][0][0];
            if (ex) acc = acc + v[q][0][1];
            if (ey) acc = acc + v[q][1][0];
            if (ex && ey) acc = acc + v[q][1][1];
            o[q] = to_store<S>(acc, bad);
        }
        *(s4 *)(dst + elem_index(slot, i, 0, C, 4)) = o;
    }
}

// Block shape of the quad relayouts (256 threads): 2^txs column pairs of a
// row pair by 256 >> txs row pairs, txs chosen on the host
// (relayout_pair_shift) so that narrow grids do not leave most lanes of a
// block's row idle.
template <typename T, typename S, bool VEC>
__global__ __launch_bounds__(256) void k_relayout_quad(const T *__restrict__ src, int W, int H, int L, int C,
                                                       LevelGeom g0, S *d00, S *d01, int coarse, LevelGeom g1, S *d10,
                                                       S *d11, int *flag, const int *only_if, int txs) {
    if (only_if && *only_if == 0) return;  // conditional redo in T: nothing to redo
    const int k = blockIdx.x * (1 << txs) + (threadIdx.x & ((1 << txs) - 1));  // column pair
    const int Y = blockIdx.y * (256 >> txs) + (threadIdx.x >> txs);           // row pair
    if (2 * k >= W || 2 * Y >= H) return;
    bool bad = false;
    for (int i = blockIdx.z; i < C; i += gridDim.z) {
        T v[4][2][2];
        relayout_quad_pair<T, S, VEC>(src, W, H, L, C, g0, d00, d01, coarse, g1, d10, d11, k, Y, i, bad, v);
    }
    report_bad(bad, flag);
}

// The same relayout over four rows per block (two row pairs), which also
// writes the warm-start network's input: the costs min-pooled over 4 x 4
// pixel blocks, times `scale`, rounded to BF16, NHWC [h4][w4][L] -- the values
// of ginf_minpool_costs, without reading the cost volume a second time.  A
// lane's two row pairs give the minimum over 2 columns x 4 rows; the
// neighbouring lane (the next column pair) completes the 4 x 4 block.
template <typename T, typename S, bool VEC>
__global__ __launch_bounds__(256) void k_relayout_quad_mp(const T *__restrict__ src, int W, int H, int L, int C,
                                                          LevelGeom g0, S *d00, S *d01, int coarse, LevelGeom g1,
                                                          S *d10, S *d11, int *flag, uint16_t *__restrict__ mp, int w4,
                                                          float scale, int txs) {
    const int k = blockIdx.x * (1 << txs) + (threadIdx.x & ((1 << txs) - 1));  // column pair
    const int Y4 = blockIdx.y * (256 >> txs) + (threadIdx.x >> txs);          // pair of row pairs
    // no early exit: lanes k and k ^ 1 (adjacent lanes, 2^txs >= 16) exchange their minima
    const bool live = 2 * k < W && 4 * Y4 < H;
    bool bad = false;
    for (int i = blockIdx.z; i < C; i += gridDim.z) {
        float mn[4] = {3.4e38f, 3.4e38f, 3.4e38f, 3.4e38f};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int Y = 2 * Y4 + h;
            if (!live || 2 * Y >= H) continue;
            T v[4][2][2];
            relayout_quad_pair<T, S, VEC>(src, W, H, L, C, g0, d00, d01, coarse, g1, d10, d11, k, Y, i, bad, v);
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                        if ((r == 0 || 2 * Y + 1 < H) && (c == 0 || 2 * k + 1 < W)) {
                            const float f = (float)v[q][r][c];
                            mn[q] = f < mn[q] ? f : mn[q];
                        }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float o = __shfl_xor(mn[q], 1);
            mn[q] = o < mn[q] ? o : mn[q];
        }
        if (live && (k & 1) == 0) {
            uint16_t *dst = mp + ((int64_t)Y4 * w4 + (k >> 1)) * L + 4 * i;
            if (4 * i + 3 < L && (L & 3) == 0) {
                *reinterpret_cast<uint2 *>(dst) = uint2{(uint32_t)f2bf(mn[0] * scale) |
                                                             ((uint32_t)f2bf(mn[1] * scale) << 16),
                                                         (uint32_t)f2bf(mn[2] * scale) |
                                                             ((uint32_t)f2bf(mn[3] * scale) << 16)};
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (4 * i + q < L) dst[q] = f2bf(mn[q] * scale);
            }
        }
    }
    report_bad(bad, flag);
}

// Data-cost pyramid level for the label-quad layout: the rule of k_pyramid,
// one thread per (plane i, coarse node) moving the node's label quad as one
// vector per child (4 B in 8-bit storage, 8 B in 16-bit, 16 B in T),
// grid-stride.  The fine level is stored in SI, the coarse one in S.
template <typename T, typename S, typename SI = S>
__global__ __launch_bounds__(256) void k_pyramid_quad(const SI *f0, const SI *f1, LevelGeom fg, S *c0, S *c1,
                                                      LevelGeom cg, int C, int *flag, const int *only_if) {
    if (only_if && *only_if == 0) return;  // conditional redo in T: nothing to redo
    typedef S s4 __attribute__((ext_vector_type(4)));
    typedef SI si4 __attribute__((ext_vector_type(4)));
    const int64_t nodes = (int64_t)cg.w * cg.h;
    const int64_t total = nodes * C;
    bool bad = false;
    for (int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
         gid += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(gid / nodes);
        const int node = (int)(gid - (int64_t)i * nodes);
        const int X = node % cg.w, Y = node / cg.w;
        const bool ex = 2 * X + 1 < fg.w, ey = 2 * Y + 1 < fg.h;
        const int64_t e00 = elem_index((int64_t)(2 * Y) * fg.whp + X, i, 0, C, 4);
        const int64_t e01 = elem_index((int64_t)(2 * Y + 1) * fg.whp + X, i, 0, C, 4);
        const si4 v00 = *(const si4 *)(f0 + e00);
        si4 v10 = si4(0), v01 = si4(0), v11 = si4(0);
        if (ex) v10 = *(const si4 *)(f1 + e00);
        if (ey) v01 = *(const si4 *)(f1 + e01);
        if (ex && ey) v11 = *(const si4 *)(f0 + e01);
        s4 o;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T acc = T(0);
            acc = acc + from_store<SI, T>(v00[q]);
            if (ex) acc = acc + from_store<SI, T>(v10[q]);
            if (ey) acc = acc + from_store<SI, T>(v01[q]);
            if (ex && ey) acc = acc + from_store<SI, T>(v11[q]);
            o[q] = to_store<S>(acc, bad);
        }
        S *dst = ((X + Y) & 1) ? c1 : c0;
        *(s4 *)(dst + elem_index((int64_t)Y * cg.whp + (X >> 1), i, 0, C, 4)) = o;
    }
    report_bad(bad, flag);
}

}  // namespace

hipError_t launch_sweep(int dtype, const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                        hipStream_t st, bool data16) {
    if (data16 && !sh.quad) return hipErrorInvalidValue;
    if (sh.quad)
        return data16 ? launch_sweep_quad16(dtype, sh, a, zeroin, labels, st)
                      : launch_sweep_quad(dtype, sh, a, zeroin, labels, st);
    if (sh.general) return launch_sweep_general(dtype, sh, a, zeroin, labels, st);
    return dtype == 0 ? launch_sweep_chunked_f32(sh, a, zeroin, labels, st)
                      : launch_sweep_chunked_i32(sh, a, zeroin, labels, st);
}
hipError_t launch_relayout(int dtype, const void *src, int W, int H, int L, const StateLayout &ly, const LevelGeom &g,
                           void *dst0, void *dst1, hipStream_t st) {
    const dim3 grid((unsigned)((W + 255) / 256), (unsigned)H), block(256);
    if (dtype == 0)
        hipLaunchKernelGGL(k_relayout<float>, grid, block, 0, st, (const float *)src, W, H, L, ly, g, (float *)dst0,
                           (float *)dst1);
    else
        hipLaunchKernelGGL(k_relayout<int>, grid, block, 0, st, (const int *)src, W, H, L, ly, g, (int *)dst0,
                           (int *)dst1);
    return hipGetLastError();
}

// log2 of the column pairs per block of the quad relayouts: the power of two
// in [16, 256] that leaves the fewest idle lanes on a row (the larger on a
// tie), e.g. 64 for W = 1920 or 640 (15 and 5 full blocks instead of 3.75 and
// 1.25 blocks of 256).
static int relayout_pair_shift(int W) {
    const int64_t P = ((int64_t)W + 1) / 2;
    int best = 8;
    int64_t waste = -1;
    for (int s = 8; s >= 4; --s) {
        const int64_t tx = (int64_t)1 << s, w = (P + tx - 1) / tx * tx - P;
        if (waste < 0 || w < waste) { best = s; waste = w; }
    }
    return best;
}

hipError_t launch_relayout_quad_mp(int dtype, const void *src, int W, int H, int L, const StateLayout &ly,
                                   const LevelGeom &g0, void *d00, void *d01, bool coarse, const LevelGeom &g1,
                                   void *d10, void *d11, hipStream_t st, bool data16, int *flag, uint16_t *mp,
                                   float scale, bool data8) {
    if (!(ly.interleave && ly.G == 4)) return hipErrorNotSupported;
    const int txs = relayout_pair_shift(W);
    const int64_t gx = (((int64_t)W + 1) / 2 + (1 << txs) - 1) >> txs;
    const int64_t gy = (((int64_t)H + 3) / 4 + (256 >> txs) - 1) / (256 >> txs);
    if (ly.C > 65535 || gy > 65535 || gx > 0x7fffffff) return hipErrorInvalidValue;
    const dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)ly.C), block(256);
    const int w4 = (W + 3) / 4;
    const bool vec = W % 2 == 0 && (uintptr_t)src % 8 == 0;
#define GINF_RQM(TT, SS)                                                                                         \
    do if (vec)                                                                                                  \
        hipLaunchKernelGGL((k_relayout_quad_mp<TT, SS, true>), grid, block, 0, st, (const TT *)src, W, H, L,      \
                           ly.C, g0, (SS *)d00, (SS *)d01, coarse ? 1 : 0, g1, (SS *)d10, (SS *)d11, flag, mp, w4, \
                           scale, txs);                                                                          \
    else                                                                                                         \
        hipLaunchKernelGGL((k_relayout_quad_mp<TT, SS, false>), grid, block, 0, st, (const TT *)src, W, H, L,     \
                           ly.C, g0, (SS *)d00, (SS *)d01, coarse ? 1 : 0, g1, (SS *)d10, (SS *)d11, flag, mp, w4, \
                           scale, txs);                                                                          \
    while (0)
    if (data8) { if (dtype == 0) GINF_RQM(float, uint8_t); else GINF_RQM(int, uint8_t); }
    else if (data16) { if (dtype == 0) GINF_RQM(float, uint16_t); else GINF_RQM(int, uint16_t); }
    else { if (dtype == 0) GINF_RQM(float, float); else GINF_RQM(int, int); }
#undef GINF_RQM
    return hipGetLastError();
}

hipError_t launch_relayout_pyr(int dtype, const void *src, int W, int H, int L, const StateLayout &ly,
                               const LevelGeom &g0, void *d00, void *d01, bool coarse, const LevelGeom &g1, void *d10,
                               void *d11, hipStream_t st, bool data16, int *flag, const int *only_if, bool data8) {
    if (ly.interleave && ly.G == 4) {  // label quads: register-only relayout
        const int txs = relayout_pair_shift(W);
        const int64_t gx = (((int64_t)W + 1) / 2 + (1 << txs) - 1) >> txs;
        const int64_t gy = (((int64_t)H + 1) / 2 + (256 >> txs) - 1) / (256 >> txs);
        if (ly.C > 65535 || gy > 65535 || gx > 0x7fffffff) return hipErrorInvalidValue;
        const dim3 grid((unsigned)gx, (unsigned)gy, only_if ? 1u : (unsigned)ly.C), block(256);
        const bool vec = W % 2 == 0 && (uintptr_t)src % 8 == 0;
#define GINF_RQ(TT, SS)                                                                                          \
        do if (vec)                                                                                              \
            hipLaunchKernelGGL((k_relayout_quad<TT, SS, true>), grid, block, 0, st, (const TT *)src, W, H, L,     \
                               ly.C, g0, (SS *)d00, (SS *)d01, coarse ? 1 : 0, g1, (SS *)d10, (SS *)d11, flag,    \
                               only_if, txs);                                                                    \
        else                                                                                                     \
            hipLaunchKernelGGL((k_relayout_quad<TT, SS, false>), grid, block, 0, st, (const TT *)src, W, H, L,    \
                               ly.C, g0, (SS *)d00, (SS *)d01, coarse ? 1 : 0, g1, (SS *)d10, (SS *)d11, flag,    \
                               only_if, txs);                                                                    \
        while (0)
        if (data8) { if (dtype == 0) GINF_RQ(float, uint8_t); else GINF_RQ(int, uint8_t); }
        else if (data16) { if (dtype == 0) GINF_RQ(float, uint16_t); else GINF_RQ(int, uint16_t); }
        else { if (dtype == 0) GINF_RQ(float, float); else GINF_RQ(int, int); }
#undef GINF_RQ
        return hipGetLastError();
    }
    const int LP = ly.C * ly.G;
    const int tw = (LP <= GINF_RELAYOUT_TW64_MAX_LP && 32 * ly.G <= 256) ? 64 : 32;
    const size_t lds = (size_t)2 * LP * (tw + 1) * 4;
    if (lds > 160 * 1024 || (tw / 2) * ly.G > 256) {
        // Layouts past the LDS tile (32 or 64 lanes per node, or more than
        // about 600 labels): per-column relayout of level 0, then level 1
        // summed from it.  Only T storage reaches here (16- and 8-bit data
        // costs are quad-layout forms).
        if (data16 || data8 || flag || only_if) return hipErrorInvalidValue;
        const dim3 grid((unsigned)((W + 255) / 256), (unsigned)H), block(256);
        if (dtype == 0)
            hipLaunchKernelGGL(k_relayout<float>, grid, block, 0, st, (const float *)src, W, H, L, ly, g0,
                               (float *)d00, (float *)d01);
        else
            hipLaunchKernelGGL(k_relayout<int>, grid, block, 0, st, (const int *)src, W, H, L, ly, g0, (int *)d00,
                               (int *)d01);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && coarse)
            e = launch_pyramid(dtype, d00, d01, g0, d10, d11, g1, ly.C, ly.G, st, false, nullptr, nullptr);
        return e;
    }
    if (data8) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((W + tw - 1) / tw), (unsigned)((H + 1) / 2)), block(256);
#define GINF_RP(TT, SS, TWW)                                                                                     \
    hipLaunchKernelGGL((k_relayout_pyr<TT, SS, TWW>), grid, block, lds, st, (const TT *)src, W, H, L, ly, g0,      \
                       (SS *)d00, (SS *)d01, coarse ? 1 : 0, g1, (SS *)d10, (SS *)d11, flag, only_if)
    if (data16) {
        if (dtype == 0) { if (tw == 32) GINF_RP(float, uint16_t, 32); else GINF_RP(float, uint16_t, 64); }
        else { if (tw == 32) GINF_RP(int, uint16_t, 32); else GINF_RP(int, uint16_t, 64); }
    } else {
        if (dtype == 0) { if (tw == 32) GINF_RP(float, float, 32); else GINF_RP(float, float, 64); }
        else { if (tw == 32) GINF_RP(int, int, 32); else GINF_RP(int, int, 64); }
    }
#undef GINF_RP
    return hipGetLastError();
}

hipError_t launch_pyramid(int dtype, const void *f0, const void *f1, const LevelGeom &fg, void *c0, void *c1,
                          const LevelGeom &cg, int C, int G, hipStream_t st, bool data16, int *flag,
                          const int *only_if, bool src8) {
    if (src8 && !(data16 && G == 4)) return hipErrorInvalidValue;
    if (G == 4) {
        // four lanes per node: one vector per child (the sums are per element,
        // so the label order of the layout does not matter)
        const int64_t threads = (int64_t)cg.w * cg.h * C;
        int64_t blocks = (threads + 255) / 256;
        const int64_t cap = only_if ? 1024 : 16384;  // the redo usually exits at once
        if (blocks > cap) blocks = cap;
        const dim3 grid((unsigned)blocks), block(256);
#define GINF_PQ(TT, SS)                                                                                           \
        hipLaunchKernelGGL((k_pyramid_quad<TT, SS>), grid, block, 0, st, (const SS *)f0, (const SS *)f1, fg,       \
                           (SS *)c0, (SS *)c1, cg, C, flag, only_if)
        if (src8) {  // 8-bit fine level (level 1 of an 8-bit run), 16-bit coarse level
            if (dtype == 0)
                hipLaunchKernelGGL((k_pyramid_quad<float, uint16_t, uint8_t>), grid, block, 0, st, (const uint8_t *)f0,
                                   (const uint8_t *)f1, fg, (uint16_t *)c0, (uint16_t *)c1, cg, C, flag, only_if);
            else
                hipLaunchKernelGGL((k_pyramid_quad<int, uint16_t, uint8_t>), grid, block, 0, st, (const uint8_t *)f0,
                                   (const uint8_t *)f1, fg, (uint16_t *)c0, (uint16_t *)c1, cg, C, flag, only_if);
        } else if (data16) { if (dtype == 0) GINF_PQ(float, uint16_t); else GINF_PQ(int, uint16_t); }
        else { if (dtype == 0) GINF_PQ(float, float); else GINF_PQ(int, int); }
#undef GINF_PQ
        return hipGetLastError();
    }
    const int64_t threads = (int64_t)cg.w * cg.h * G * C;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
#define GINF_PY(TT, SS)                                                                                            \
    hipLaunchKernelGGL((k_pyramid<TT, SS>), grid, block, 0, st, (const SS *)f0, (const SS *)f1, fg, (SS *)c0,      \
                       (SS *)c1, cg, C, G, flag, only_if)
    if (data16) { if (dtype == 0) GINF_PY(float, uint16_t); else GINF_PY(int, uint16_t); }
    else { if (dtype == 0) GINF_PY(float, float); else GINF_PY(int, int); }
#undef GINF_PY
    return hipGetLastError();
}

hipError_t launch_upsample(int dtype, const UpsampleArgs &u, hipStream_t st) {
    const int64_t threads = (int64_t)u.fg.w * u.fg.h * u.G;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    if (dtype == 0) hipLaunchKernelGGL(k_upsample<float>, grid, block, 0, st, u);
    else hipLaunchKernelGGL(k_upsample<int>, grid, block, 0, st, u);
    return hipGetLastError();
}

hipError_t launch_fill_u32(void *p, int64_t n, uint32_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t *)p, n, v);
    return hipGetLastError();
}

hipError_t launch_zero_labels(int32_t *labels, int W, int H, int color, hipStream_t st) {
    const dim3 grid((unsigned)((W + 255) / 256), (unsigned)H), block(256);
    hipLaunchKernelGGL(k_zero_labels, grid, block, 0, st, labels, W, H, color);
    return hipGetLastError();
}

hipError_t launch_energy(int dtype, const void *data, const void *table, const int32_t *labels, int W, int H, int L,
                         double *partials, int nblocks, double *out, hipStream_t st) {
    if (dtype == 0)
        hipLaunchKernelGGL(k_energy_partial<float>, dim3(nblocks), dim3(256), 0, st, (const float *)data,
                           (const float *)table, labels, W, H, L, partials);
    else
        hipLaunchKernelGGL(k_energy_partial<int>, dim3(nblocks), dim3(256), 0, st, (const int *)data,
                           (const int *)table, labels, W, H, L, partials);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, partials, nblocks, out);
    return hipGetLastError();
}

}  // namespace ginf
