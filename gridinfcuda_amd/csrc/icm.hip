// Iterated conditional modes and the max-prior labelling on gfx950.
//
// ICM (reference src/icm.cc:5-62, src/gpu/icm.cu:21-83): every interior node
// moves to the lowest-index label of minimal local cost
//   D(x,y,l) + V(l,f_S) + V(l,f_E) + V(l,f_N) + V(l,f_W)
// accumulated left to right in T, as Grid::getLabelingCost(f,x,y,l)
// (src/grid.cc:127-136).
//   * synchronous: all nodes read the previous iteration's labelling (two
//     label buffers, one launch per iteration);
//   * asynchronous: the reference updates in place in row-major order, so
//     node (x,y) sees this iteration's labels at (x-1,y) and (x,y-1) and the
//     previous iteration's at (x+1,y) and (x,y+1).  The nodes of one
//     anti-diagonal x+y = k are therefore independent, and iteration t+1 may
//     work on diagonal k while iteration t works on diagonal k+2: launch s
//     updates diagonal s-2t for every iteration t at once, which reproduces
//     the sequential result exactly.
// Synchronous ICM reads the costs in the reference layout [l][y][x] (a wave
// reads adjacent x of one label plane per load); asynchronous ICM first makes
// a node-major copy [y][x][l], since its launches visit one anti-diagonal of
// nodes at a time.  The L x L table V(a,b) at a + L*b is staged in LDS for
// synchronous ICM with L <= 128 and read through the cache otherwise.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "bp_kernels.h"
#include "ginf_bp.h"

namespace {

// Stream-ordered scratch allocations (label buffer, 8-bit cost copy,
// node-major cost copy) come from a memory pool of the library's own per
// device, which keeps freed memory for the next call (release threshold at
// its maximum) instead of unmapping it at every synchronisation; the process's
// default pool is left as it is.  Pool memory lives until the process ends.
// Returns nullptr when the pool cannot be created (the caller then uses the
// default pool).
hipMemPool_t scratch_pool(int dev) {
    constexpr int kMaxDev = 64;
    static hipMemPool_t pools[kMaxDev] = {};
    static std::mutex mu;
    if (dev < 0 || dev >= kMaxDev) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!pools[dev]) {
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &props) != hipSuccess) return nullptr;
        uint64_t keep = UINT64_MAX;
        if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
            (void)hipMemPoolDestroy(pool);
            return nullptr;
        }
        pools[dev] = pool;
    }
    return pools[dev];
}

hipError_t scratch_alloc(void **p, size_t n, hipStream_t st) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    hipMemPool_t pool = scratch_pool(dev);
    return pool ? hipMallocFromPoolAsync(p, n, pool, st) : hipMallocAsync(p, n, st);
}

// Neighbour labels outside [0, L) (undefined in the reference) are clamped so
// that the table reads stay inside the table.
__device__ __forceinline__ int clamp_label(int f, int L) { return f < 0 ? 0 : (f >= L ? L - 1 : f); }

template <typename T>
__device__ __forceinline__ int icm_best(const T *__restrict__ data, const T *__restrict__ table, int64_t plane,
                                        int64_t node, int L, int fS, int fE, int fN, int fW) {
    fS = clamp_label(fS, L);
    fE = clamp_label(fE, L);
    fN = clamp_label(fN, L);
    fW = clamp_label(fW, L);
    const T *vS = table + (int64_t)L * fS;
    const T *vE = table + (int64_t)L * fE;
    const T *vN = table + (int64_t)L * fN;
    const T *vW = table + (int64_t)L * fW;
    // Labels in groups of U: the group's cost loads are issued together (one
    // cost plane apart each, independent of the running minimum), then the
    // sums and the comparisons run in label order, so each wave waits for
    // memory once per group instead of once per label.
    constexpr int U = 16;
    T best = T(0);
    int arg = 0;
    for (int l0 = 0; l0 < L; l0 += U) {
        T dv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) dv[u] = l0 + u < L ? data[(int64_t)(l0 + u) * plane + node] : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int l = l0 + u;
            if (l < L) {
                T c = dv[u];
                c = c + vS[l];
                c = c + vE[l];
                c = c + vN[l];
                c = c + vW[l];
                if (l == 0 || c < best) { best = c; arg = l; }
            }
        }
    }
    return arg;
}

// ---- table in LDS (synchronous ICM, L <= kLdsMaxLabels) -----------------
// With the table in global memory every label of every node gathers V(l, f)
// for four neighbour labels f that differ from lane to lane, 64 cache lines
// per wave and load.  The kernels below stage the table in LDS once per
// block: row f (a neighbour label) holds V(l, f) for l in [0, L) at
// l ^ (f & 31) of a row of Lp = roundup(L, 32) entries, so lanes reading the
// same l from different rows fall in different banks.  The sums and the
// comparisons are those of icm_best, in the same order.
constexpr int kLdsMaxLabels = 128;  // L * Lp * 4 B <= 64 KiB

template <typename T>
__device__ __forceinline__ void stage_table(T *lds, const T *__restrict__ table, int L, int Lp) {
    for (int i = threadIdx.x; i < L * L; i += blockDim.x) {
        const int l = i % L, f = i / L;  // table[l + L * f] = V(l, f)
        lds[f * Lp + (l ^ (f & 31))] = table[i];
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ int icm_best_lds(const T *__restrict__ data, const T *tab, int Lp, int64_t plane,
                                            int64_t node, int L, int fS, int fE, int fN, int fW) {
    fS = clamp_label(fS, L);
    fE = clamp_label(fE, L);
    fN = clamp_label(fN, L);
    fW = clamp_label(fW, L);
    const T *rS = tab + fS * Lp, *rE = tab + fE * Lp, *rN = tab + fN * Lp, *rW = tab + fW * Lp;
    const int sS = fS & 31, sE = fE & 31, sN = fN & 31, sW = fW & 31;
    constexpr int U = 16;
    T best = T(0);
    int arg = 0;
    for (int l0 = 0; l0 < L; l0 += U) {
        T dv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) dv[u] = l0 + u < L ? data[(int64_t)(l0 + u) * plane + node] : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int l = l0 + u;
            if (l < L) {
                T c = dv[u];
                c = c + rS[l ^ sS];
                c = c + rE[l ^ sE];
                c = c + rN[l ^ sN];
                c = c + rW[l ^ sW];
                if (l == 0 || c < best) { best = c; arg = l; }
            }
        }
    }
    return arg;
}

// Synchronous iteration over the interior as one flat index (x fastest),
// grid-stride with a grid sized to the machine, so each block stages the
// table once.
template <typename T>
__global__ __launch_bounds__(1024) void k_icm_sync_lds(const T *__restrict__ data, const T *__restrict__ table, int W,
                                                       int H, int L, const int32_t *__restrict__ old,
                                                       int32_t *__restrict__ out) {
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);
    const int Lp = (L + 31) & ~31;
    stage_table(tab, table, L, Lp);
    const int wi = W - 2;
    const int64_t nin = (int64_t)wi * (H - 2);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nin; i += (int64_t)gridDim.x * blockDim.x) {
        const int y = 1 + (int)(i / wi), x = 1 + (int)(i - (int64_t)(y - 1) * wi);
        const int64_t n = (int64_t)y * W + x;
        out[n] = icm_best_lds(data, tab, Lp, (int64_t)W * H, n, L, old[n + W], old[n + 1], old[n - W], old[n - 1]);
    }
}

// The same iteration with four nodes x0..x0+3 (x0 % 4 == 0) per thread, for
// W % 4 == 0, L % 4 == 0 and 16-byte aligned costs and label buffers: one
// 16-byte load brings a label's costs of the four nodes (a wave reads 1 KiB
// of one plane per load instead of 256 B), and the labels are read and
// written as int4.  A quad may hold the border columns x = 0 and x = W-1;
// those nodes keep their old label, which both label buffers hold unchanged.
//
// The table is staged in label quads: row f holds V(l, f) at
// l ^ ((f & 7) << 2) of a row of Lr = max(32, 2^ceil(log2 L)) entries, so the
// four labels l0..l0+3 (l0 % 4 == 0) of a row are one aligned 16-byte LDS
// read, rows with different f & 7 start their quads in different banks, and
// with the power-of-two stride a quad's byte address is (row base | swizzle)
// XOR 4*l0: one VALU operation per neighbour and four labels.  Per node the
// sums and the comparisons are those of icm_best, in the same order.
template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<int> { using type = int4; };

__host__ __device__ inline int quad_row_stride(int L) {
    int r = 32;
    while (r < L) r <<= 1;
    return r;
}

template <typename T>
__device__ __forceinline__ void stage_table_quads(T *lds, const T *__restrict__ table, int L, int Lr) {
    for (int i = threadIdx.x; i < L * L; i += blockDim.x) {
        const int l = i % L, f = i / L;  // table[l + L * f] = V(l, f)
        lds[f * Lr + (l ^ ((f & 7) << 2))] = table[i];
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ T lane_of(const typename Vec4<T>::type &v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}

// Four nodes' costs of one label: a 16-byte load of T, or a 4-byte load of
// the 8-bit copy (exact: every cost is an integer in [0, 255]).
template <typename T, typename S>
__device__ __forceinline__ void load_cost4(const S *p, T (&c)[4]) {
    if constexpr (sizeof(S) == 1) {
        const uint32_t v = *reinterpret_cast<const uint32_t *>(p);
        c[0] = T(v & 0xFFu);
        c[1] = T((v >> 8) & 0xFFu);
        c[2] = T((v >> 16) & 0xFFu);
        c[3] = T(v >> 24);
    } else {
        const typename Vec4<T>::type v = *reinterpret_cast<const typename Vec4<T>::type *>(p);
        c[0] = v.x;
        c[1] = v.y;
        c[2] = v.z;
        c[3] = v.w;
    }
}

// Exact 8-bit form of a cost: an integer in [0, 255] (and not -0).
template <typename T>
__device__ __forceinline__ bool fits_u8(T e) {
    if constexpr (std::is_same<T, float>::value)
        return e >= 0.f && e <= 255.f && e == __builtin_truncf(e) && !__builtin_signbit(e);
    else
        return e >= 0 && e <= 255;
}

// One synchronous iteration over the interior, four nodes per thread
// (grid-stride), the table already in LDS.  S = T reads the costs in T; S =
// uint8_t reads their 8-bit copy.  COPY (with S = T): also writes the 8-bit
// copy of every cost it reads into `copy` (same [l][y][x] layout) and sets
// *bad for a cost with no exact 8-bit form.
template <typename T, typename S, bool COPY>
__device__ __forceinline__ void icm_x4_pass(const S *__restrict__ data, const unsigned char *tab_raw, int W, int H, int L,
                                            const int32_t *__restrict__ old, int32_t *__restrict__ out,
                                            uint8_t *__restrict__ copy, bool &bad) {
    using V4 = typename Vec4<T>::type;
    const int Lr = quad_row_stride(L);
    const int qw = W >> 2;
    const int64_t nq = (int64_t)qw * (H - 2), plane = (int64_t)W * H;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += (int64_t)gridDim.x * blockDim.x) {
        const int y = 1 + (int)(i / qw), x0 = 4 * (int)(i - (int64_t)(y - 1) * qw);
        const int64_t n0 = (int64_t)y * W + x0;
        const int4 fc = *reinterpret_cast<const int4 *>(old + n0);
        const int4 fs = *reinterpret_cast<const int4 *>(old + n0 + W);
        const int4 fn = *reinterpret_cast<const int4 *>(old + n0 - W);
        const int fw0 = x0 > 0 ? old[n0 - 1] : 0;      // only read for an interior node
        const int fe3 = x0 + 4 < W ? old[n0 + 4] : 0;  // only read for an interior node
        const int cur[4] = {fc.x, fc.y, fc.z, fc.w};
        // S, E, N, W neighbour labels of node k
        const int nb[4][4] = {{fs.x, fc.y, fn.x, fw0}, {fs.y, fc.z, fn.y, fc.x},
                              {fs.z, fc.w, fn.z, fc.y}, {fs.w, fe3, fn.w, fc.z}};
        int P[4][4];  // byte address of row f's quad 0, swizzle included
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int f = clamp_label(nb[k][d], L);
                P[k][d] = (f * Lr * (int)sizeof(T)) | ((f & 7) << 4);
            }
        T best[4] = {T(0), T(0), T(0), T(0)};
        int arg[4] = {0, 0, 0, 0};
        // Labels in quads, software-pipelined: the next quad's cost loads are
        // issued before the current quad's sums, so every wave keeps loads in
        // flight while it computes (the last quad re-reads plane L-1 instead
        // of branching).
        T dv[4][4], nx[4][4];  // [label l0 + u][node k]
#pragma unroll
        for (int u = 0; u < 4; ++u) load_cost4(data + (int64_t)u * plane + n0, dv[u]);
        for (int l0 = 0; l0 < L; l0 += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) load_cost4(data + (int64_t)min(l0 + 4 + u, L - 1) * plane + n0, nx[u]);
            if constexpr (COPY) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    uint32_t o = 0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const bool ok = fits_u8(dv[u][k]);
                        bad |= !ok;
                        o |= (ok ? (uint32_t)dv[u][k] : 0u) << (8 * k);
                    }
                    *reinterpret_cast<uint32_t *>(copy + (int64_t)(l0 + u) * plane + n0) = o;
                }
            }
            const int lb = l0 * (int)sizeof(T);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                V4 t[4];  // V(l0..l0+3, f) for the S, E, N, W neighbours of node k
#pragma unroll
                for (int d = 0; d < 4; ++d) t[d] = *reinterpret_cast<const V4 *>(tab_raw + (P[k][d] ^ lb));
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int l = l0 + u;
                    T c = dv[u][k];
                    c = c + lane_of<T>(t[0], u);
                    c = c + lane_of<T>(t[1], u);
                    c = c + lane_of<T>(t[2], u);
                    c = c + lane_of<T>(t[3], u);
                    if (l == 0 || c < best[k]) { best[k] = c; arg[k] = l; }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int k = 0; k < 4; ++k) dv[u][k] = nx[u][k];
        }
        const bool xl = x0 == 0, xr = x0 + 4 == W;  // border columns keep their label
        *reinterpret_cast<int4 *>(out + n0) =
            int4{xl ? cur[0] : arg[0], arg[1], arg[2], xr ? cur[3] : arg[3]};
    }
}

// Synchronous iteration with four nodes per thread.  MODE 0 reads the costs
// in T; MODE 1 (the first iteration of a call) reads them in T and writes
// their 8-bit copy `c8` and *flag (set when a cost has no exact 8-bit form);
// MODE 2 (the later iterations) reads the copy when *flag is clear, else the
// costs in T -- a quarter of the cost bytes per iteration for stereo-style
// costs, decided on the device.  The copy holds every node of every quad an
// iteration reads (rows 1..H-2); a border column's costs are read and
// discarded.
template <typename T, int MODE>
__global__ __launch_bounds__(768) void k_icm_sync_lds_x4(const T *__restrict__ data, uint8_t *__restrict__ c8,
                                                          int *__restrict__ flag, const T *__restrict__ table, int W,
                                                          int H, int L, const int32_t *__restrict__ old,
                                                          int32_t *__restrict__ out) {
    extern __shared__ unsigned char smem_raw[];
    stage_table_quads(reinterpret_cast<T *>(smem_raw), table, L, quad_row_stride(L));
    bool bad = false;
    if constexpr (MODE == 2) {
        if (__builtin_amdgcn_readfirstlane(*flag) == 0)
            icm_x4_pass<T, uint8_t, false>(c8, smem_raw, W, H, L, old, out, nullptr, bad);
        else
            icm_x4_pass<T, T, false>(data, smem_raw, W, H, L, old, out, nullptr, bad);
    } else if constexpr (MODE == 1) {
        icm_x4_pass<T, T, true>(data, smem_raw, W, H, L, old, out, c8, bad);
        if (bad) atomicOr(flag, 1);
    } else {
        icm_x4_pass<T, T, false>(data, smem_raw, W, H, L, old, out, nullptr, bad);
    }
}

// Launch s of asynchronous ICM: iteration t = t0 + blockIdx.y updates
// anti-diagonal j = s - 2t (x + y = 2 + j) of the interior, in place, with 16
// lanes per node: lane q takes labels q, q+16,
// ... (so the 16 lanes of a node read 16 consecutive entries of each table
// column), forms each label's cost in the order of icm_best, keeps its
// lowest-index minimum, and the lanes combine: the smaller cost wins, ties go
// to the lower label.  That is the sequential scan's result for finite costs,
// and a step's few nodes now wait for L/16 table reads each instead of L.
constexpr int kAsyncLanes = 16;
constexpr int kAsyncChunk = 8;  // labels per lane per round of loads (L <= 128: one round)

// Costs come from the node-major copy made by k_costs_node_major (element
// (l, n) at n * L + l), so the 16 lanes of a node read 64 contiguous bytes
// per step; in the reference layout every node of a diagonal touched its own
// cache line in every plane (1920x1080x128, 10 iterations: 33.5 ms with the
// reference layout, 18.9 ms with the copy, whose one pass costs ~0.5 ms, and
// 14.5 ms with each chunk's loads issued together).
template <typename T>
__global__ __launch_bounds__(256) void k_icm_async_lanes(const T *__restrict__ data, const T *__restrict__ table,
                                                         int W, int H, int L, int32_t *__restrict__ lab, int s, int t0,
                                                         int nd) {
    const int t = t0 + blockIdx.y;
    const int j = s - 2 * t;
    if (j < 0 || j >= nd) return;  // uniform per block
    const int k = 2 + j;
    const int xs = max(1, k - (H - 2)), xe = min(W - 2, k - 1);
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = xs + gid / kAsyncLanes, q = gid % kAsyncLanes;
    const bool active = x <= xe;  // inactive lanes still join the shuffles
    const int y = k - x;
    const int64_t n = (int64_t)y * W + x;
    T best = T(0);
    int arg = -1;
    if (active) {
        const int fS = clamp_label(lab[n + W], L), fE = clamp_label(lab[n + 1], L);
        const int fN = clamp_label(lab[n - W], L), fW = clamp_label(lab[n - 1], L);
        const T *vS = table + (int64_t)L * fS, *vE = table + (int64_t)L * fE;
        const T *vN = table + (int64_t)L * fN, *vW = table + (int64_t)L * fW;
        // Labels q, q+16, ... in chunks of kAsyncChunk whose cost and table
        // loads are all issued before the chunk's sums (indices past L - 1
        // re-read label L - 1 and are not compared).
        for (int l0 = q; l0 < L; l0 += kAsyncChunk * kAsyncLanes) {
            T dv[kAsyncChunk], tS[kAsyncChunk], tE[kAsyncChunk], tN[kAsyncChunk], tW[kAsyncChunk];
#pragma unroll
            for (int u = 0; u < kAsyncChunk; ++u) {
                const int l = min(l0 + u * kAsyncLanes, L - 1);
                dv[u] = data[n * L + l];
                tS[u] = vS[l];
                tE[u] = vE[l];
                tN[u] = vN[l];
                tW[u] = vW[l];
            }
#pragma unroll
            for (int u = 0; u < kAsyncChunk; ++u) {
                const int l = l0 + u * kAsyncLanes;
                if (l < L) {
                    T c = dv[u];
                    c = c + tS[u];
                    c = c + tE[u];
                    c = c + tN[u];
                    c = c + tW[u];
                    if (arg < 0 || c < best) { best = c; arg = l; }
                }
            }
        }
    }
#pragma unroll
    for (int m = 1; m < kAsyncLanes; m <<= 1) {
        const T ob = __shfl_xor(best, m, kAsyncLanes);
        const int oa = __shfl_xor(arg, m, kAsyncLanes);
        if (oa >= 0 && (arg < 0 || ob < best || (!(best < ob) && oa < arg))) { best = ob; arg = oa; }
    }
    if (active && q == 0) lab[n] = arg;
}

// ---- asynchronous ICM on tiles (L <= kLdsMaxLabels) -----------------------
// The row-major in-place order is a wavefront over anti-diagonals; here the
// wavefront runs over square tiles of TS x TS interior nodes.  Launch s
// updates, for every iteration t with 0 <= d = s - 2t < tile diagonals, the
// tiles (tx, ty) with tx + ty = d: a tile needs this iteration's labels of
// its west and north tiles (diagonal d - 1, earlier launches) and the
// previous iteration's of its east and south tiles (diagonal d + 1 of
// iteration t - 1, launch s - 1; their iteration t comes at launch s + 1),
// and no two tiles of one launch touch each other.  A block stages the
// table in LDS (as synchronous ICM) and its tile's labels with a one-node
// halo, then walks the tile's node anti-diagonals in order with a barrier
// after each, 16 lanes per node as k_icm_async_lanes (same sums, same
// tie-breaking), and writes the tile back.  The costs (node-major copy) of
// the diagonal kPrefetch steps ahead are loaded while a step runs.  One
// launch per tile diagonal replaces one per node diagonal.
#ifndef GINF_ICM_TILE
#define GINF_ICM_TILE 32
#endif

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// operations and meets the other waves, without the release/acquire fence of
// __syncthreads(), which also waits for every global load in flight (here the
// cost prefetches of the next steps).  The memory clobber keeps the compiler
// from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
#ifndef GINF_ICM_TILE_LANES
#define GINF_ICM_TILE_LANES 16
#endif
constexpr int kTileLanes = GINF_ICM_TILE_LANES;  // lanes per node (8 or 16)
static_assert(kTileLanes == 8 || kTileLanes == 16, "8 or 16 lanes per node");
constexpr int kPrefetch = 4;
// Tile steps per iteration of the step loop (a multiple of kPrefetch): the
// compiler waits for every load in flight at the loop's head, so longer
// iterations drain the prefetch less often.
#ifndef GINF_ICM_STEPS_PER_ITER
#define GINF_ICM_STEPS_PER_ITER 16
#endif
constexpr int kStepsPerIter = GINF_ICM_STEPS_PER_ITER;
static_assert(kStepsPerIter % kPrefetch == 0, "steps per loop iteration: a multiple of the prefetch depth");

// stage_table with every load of a thread issued before its LDS stores (one
// memory round trip per block instead of one per element group), over a
// 128-column index space (row f = idx / 128, label l = idx % 128, no division
// by L): the tile kernel stages the table once per launch, so its staging
// time counts.
template <typename T, int NT>
__device__ __forceinline__ void stage_table_batched(T *lds, const T *__restrict__ table, int L, int Lp) {
    constexpr int PER = (kLdsMaxLabels * kLdsMaxLabels + NT - 1) / NT;
    T v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {  // unconditional loads (clamped in-bounds), so they issue back to back
        const int idx = threadIdx.x + k * NT, f = idx / kLdsMaxLabels, l = idx % kLdsMaxLabels;
        v[k] = table[min(f, L - 1) * L + min(l, L - 1)];  // table[l + L * f] = V(l, f)
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int idx = threadIdx.x + k * NT, f = idx / kLdsMaxLabels, l = idx % kLdsMaxLabels;
        if (f < L && l < L) lds[f * Lp + (l ^ (f & 31))] = v[k];
    }
    __syncthreads();
}

// (cost, label) pair of a lane's lowest-index minimum; combine keeps the
// smaller cost, on a tie the lower label (-1: no label).  For costs that are
// not NaN the rule is commutative and associative, so any reduction order
// gives the same pair: the sequential scan's.
template <typename T>
__device__ __forceinline__ void icm_combine(T &best, int &arg, T ob, int oa) {
    if (oa >= 0 && (arg < 0 || ob < best || (!(best < ob) && oa < arg))) { best = ob; arg = oa; }
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_mov(T v) {
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

template <typename T, int TS, int NL>
__global__ __launch_bounds__(TS * NL) void k_icm_async_tiles(const T *__restrict__ nm,
                                                                     const T *__restrict__ table, int W, int H, int L,
                                                                     int32_t *__restrict__ lab, int s, int t0,
                                                                     int ntx, int nty, uint32_t nm_bytes) {
    extern __shared__ __attribute__((aligned(16))) unsigned char icm_smem[];
    const int t = t0 + blockIdx.y;
    const int d = s - 2 * t;
    if (d < 0 || d >= ntx + nty - 1) return;  // uniform per block
    const int tx = max(0, d - (nty - 1)) + blockIdx.x;
    if (tx > min(ntx - 1, d)) return;
    const int ty = d - tx;
    const int x0 = 1 + tx * TS, y0 = 1 + ty * TS;
    const int tw = min(TS, W - 1 - x0), th = min(TS, H - 1 - y0);  // tile extent inside the interior
    constexpr int P = TS + 2;
    constexpr int Lp = kLdsMaxLabels;  // table rows of 128 entries: l ^ (f & 31) stays in its row for any l < 128
    __shared__ int ls[P * P];          // the tile's labels with a one-node halo
    T *tab = (T *)icm_smem;            // the table, L rows of Lp (<= 64 KiB)
    {
        // the tile's labels and halo: unconditional loads (clamped in-bounds) issued together
        constexpr int PL = (P * P + TS * NL - 1) / (TS * NL);
        int v[PL];
#pragma unroll
        for (int k = 0; k < PL; ++k) {
            const int i = threadIdx.x + k * TS * NL;
            const int gy = min(y0 - 1 + i / P, H - 1), gx = min(x0 - 1 + i % P, W - 1);
            v[k] = lab[(int64_t)gy * W + gx];
        }
#pragma unroll
        for (int k = 0; k < PL; ++k) {
            const int i = threadIdx.x + k * TS * NL;
            if (i < P * P) ls[i] = clamp_label(v[k], L);  // entries past the image are never read
        }
    }
    stage_table_batched<T, TS * NL>(tab, table, L, Lp);  // ends with a barrier
    constexpr int PER = kLdsMaxLabels / NL;  // labels per lane
    const int i = threadIdx.x / NL, q = threadIdx.x % NL;
    const int nsteps = tw + th - 1;
    // node of this thread at tile step k (lx < 0: none)
    auto node_at = [&](int k, int &lx, int &ly) {
        const int lo = max(0, k - (th - 1)), hi = min(tw - 1, k);
        lx = lo + i;
        ly = k - lx;
        if (k >= nsteps || lx > hi) lx = -1;
    };
    // Costs through a buffer resource with 32-bit offsets: a lane's labels
    // q, q + NL, ... of a node are NL * sizeof(T) bytes apart (immediate offsets);
    // labels past L - 1 read the next node's costs (or 0 past the copy's end)
    // and are not compared.
    const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc((void *)nm, 0, nm_bytes, 0x00020000);
    // Straight-line steps (no branch around a load, every prefetch issued,
    // steps padded to a multiple of kStepsPerIter), so that the compiler can
    // count the loads in flight and wait only for the slot a step uses.
    T buf[kPrefetch][PER];
    auto load = [&](int k, T (&dv)[PER]) {
        int lx, ly;
        node_at(k, lx, ly);
        if (lx < 0) lx = ly = 0;  // no node: a harmless load of the tile's first node
        const uint32_t off = ((uint32_t)((y0 + ly) * W + (x0 + lx)) * (uint32_t)L + (uint32_t)q) * (uint32_t)sizeof(T);
#pragma unroll
        for (int u = 0; u < PER; ++u)
            dv[u] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(crs, off + u * NL * sizeof(T), 0, 0));
    };
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) load(u, buf[u]);
    for (int k0 = 0; k0 < nsteps; k0 += kStepsPerIter) {
#pragma unroll
        for (int uu = 0; uu < kStepsPerIter; ++uu) {
            const int k = k0 + uu, u = uu % kPrefetch;
            int lx, ly;
            node_at(k, lx, ly);
            const bool act = lx >= 0;
            const int c0 = act ? (ly + 1) * P + (lx + 1) : P + 1;
            // waves without a node on this diagonal skip the step's work (a
            // uniform branch with no loads inside; the prefetch below is issued
            // by every wave)
            const int span = min(tw - 1, k) - max(0, k - (th - 1));  // nodes on the diagonal - 1
            const int wave_first = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64)) * (64 / NL);
            T best = T(0);
            int arg = -1;
            if (k < nsteps && wave_first <= span) {
                const int f4[4] = {ls[c0 + P], ls[c0 + 1], ls[c0 - P], ls[c0 - 1]};  // S, E, N, W
                // row f, label l = q + NL v sits at f * Lp + (l ^ (f & 31)) =
                // f * Lp + (q ^ (f & (NL - 1))) + NL * (v ^ sb), sb = (f & 31) / NL:
                // one base per row and value of v mod G (the low bits sb can
                // flip), the rest immediate offsets
                constexpr int G = 32 / NL;
                const T *base[4][G];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int f = f4[r];
                    const T *row = tab + f * Lp + (q ^ (f & (NL - 1)));
                    const int sb = (f & 31) / NL;
#pragma unroll
                    for (int j = 0; j < G; ++j) base[r][j] = row + NL * ((j ^ sb) - j);
                }
#pragma unroll
                for (int v = 0; v < PER; ++v) {
                    const int l = q + v * NL;
                    T c = buf[u][v];
#pragma unroll
                    for (int r = 0; r < 4; ++r) c = c + base[r][v % G][NL * v];
                    const bool take = act && l < L && (arg < 0 || c < best);
                    best = take ? c : best;
                    arg = take ? l : arg;
                }
                // reduce over the node's NL lanes: lanes 1 and 2 of each quad,
                // the two quads of each half row, (NL = 16) the two halves of the row
                icm_combine(best, arg, dpp_mov<0xB1>(best), dpp_mov<0xB1>(arg));    // quad_perm [1,0,3,2]
                icm_combine(best, arg, dpp_mov<0x4E>(best), dpp_mov<0x4E>(arg));    // quad_perm [2,3,0,1]
                icm_combine(best, arg, dpp_mov<0x141>(best), dpp_mov<0x141>(arg));  // row_half_mirror
                if constexpr (NL == 16) icm_combine(best, arg, dpp_mov<0x140>(best), dpp_mov<0x140>(arg));  // row_mirror
                if (act && q == 0) ls[c0] = arg;
            }
            load(k + kPrefetch, buf[u]);  // this slot's next diagonal
            lds_barrier();
        }
    }
    for (int j = threadIdx.x; j < tw * th; j += blockDim.x) {
        const int ly = j / tw, lx = j % tw;
        lab[(int64_t)(y0 + ly) * W + (x0 + lx)] = ls[(ly + 1) * P + (lx + 1)];
    }
}

// Node-major copy of the costs for asynchronous ICM: out[n * L + l] =
// data[l * P + n] (P = W * H), through a 32-label x 64-node LDS tile so that
// both the plane reads and the node-row writes are contiguous.
template <typename T>
__global__ __launch_bounds__(256) void k_costs_node_major(const T *__restrict__ data, int64_t P, int L,
                                                          T *__restrict__ out) {
    __shared__ T tile[32][65];
    const int64_t n0 = (int64_t)blockIdx.x * 64;
    const int l0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    for (int r = ty; r < 32; r += 4) {
        const int l = l0 + r;
        const int64_t n = n0 + tx;
        if (l < L && n < P) tile[r][tx] = data[(int64_t)l * P + n];
    }
    __syncthreads();
    const int lx = threadIdx.x & 31, ny = threadIdx.x >> 5;  // 32 x 8
    for (int c = ny; c < 64; c += 8) {
        const int64_t n = n0 + c;
        const int l = l0 + lx;
        if (l < L && n < P) out[n * L + l] = tile[lx][c];
    }
}

// ---- table in global memory (any L) ---------------------------------------
template <typename T>
__global__ void k_icm_sync(const T *__restrict__ data, const T *__restrict__ table, int W, int H, int L,
                           const int32_t *__restrict__ old, int32_t *__restrict__ out) {
    const int x = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = 1 + blockIdx.y;
    if (x > W - 2 || y > H - 2) return;
    const int64_t n = (int64_t)y * W + x;
    out[n] = icm_best(data, table, (int64_t)W * H, n, L, old[n + W], old[n + 1], old[n - W], old[n - 1]);
}

template <typename T>
__global__ void k_max_prior(const T *__restrict__ data, int W, int H, int L, int32_t *__restrict__ out) {
    const int64_t n = (int64_t)W * H;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T best = data[i];
        int arg = 0;
        for (int l = 1; l < L; ++l) {
            const T c = data[(int64_t)l * n + i];
            if (c < best) { best = c; arg = l; }
        }
        out[i] = arg;
    }
}

// Four nodes per thread (W * H % 4 == 0, 16-byte aligned buffers): one
// 16-byte load per label brings four nodes' costs, and a group of eight
// labels' loads is issued before their comparisons (labels past L - 1
// re-read label L - 1 and are not compared).  Per node the scan is that of
// k_max_prior: lowest-index label of minimal cost.
template <typename T>
__global__ __launch_bounds__(256) void k_max_prior_x4(const T *__restrict__ data, int64_t n, int L,
                                                      int32_t *__restrict__ out) {
    using V4 = typename Vec4<T>::type;
    constexpr int U = 8;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < (n >> 2); q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = 4 * q;
        const V4 b0 = *reinterpret_cast<const V4 *>(data + i);
        T best[4] = {b0.x, b0.y, b0.z, b0.w};
        int arg[4] = {0, 0, 0, 0};
        for (int l0 = 1; l0 < L; l0 += U) {
            V4 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = *reinterpret_cast<const V4 *>(data + (int64_t)min(l0 + u, L - 1) * n + i);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int l = l0 + u;
                if (l < L) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const T c = lane_of<T>(v[u], k);
                        if (c < best[k]) { best[k] = c; arg[k] = l; }
                    }
                }
            }
        }
        *reinterpret_cast<int4 *>(out + i) = int4{arg[0], arg[1], arg[2], arg[3]};
    }
}

template <typename T>
hipError_t icm_t(const T *data, const T *table, int W, int H, int L, const int32_t *initial, int iters, int mode,
                 int32_t *result, hipStream_t st) {
    const size_t nl = (size_t)W * H * sizeof(int32_t);
    hipError_t e = hipSuccess;
    if (result != initial) e = hipMemcpyAsync(result, initial, nl, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess || iters == 0 || W < 3 || H < 3) return e;
    const bool lds = L <= kLdsMaxLabels;
    const size_t lds_bytes = (size_t)L * ((L + 31) & ~31) * sizeof(T);
    const size_t lds_bytes_x4 = (size_t)L * quad_row_stride(L) * sizeof(T);
    if (mode == GINF_ICM_SYNC) {
        int32_t *tmp = nullptr;
        if ((e = scratch_alloc((void **)&tmp, nl, st)) != hipSuccess) return e;
        e = hipMemcpyAsync(tmp, initial, nl, hipMemcpyDeviceToDevice, st);
        // The last iteration writes `result`.
        int32_t *bufs[2] = {(iters & 1) ? tmp : result, (iters & 1) ? result : tmp};
        int dev = 0, cus = 0;
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const int64_t nin = (int64_t)(W - 2) * (H - 2);
        // Four nodes per thread where the quads are whole and 16-byte aligned.
        const bool x4 = lds && W % 4 == 0 && L % 4 == 0 && ((uintptr_t)data & 15) == 0 && ((uintptr_t)result & 15) == 0 &&
                        ((uintptr_t)tmp & 15) == 0;
        // (x4: 768-thread blocks, two per CU with 64 KiB of LDS each: six waves
        // per SIMD, which its 80 VGPRs allow)
        const int64_t items = x4 ? (int64_t)(W / 4) * (H - 2) : nin;
        const int bs = x4 ? 768 : 1024;
        const unsigned lds_grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((items + bs - 1) / bs, 2 * (int64_t)cus));
        const dim3 grid((unsigned)((W - 2 + 255) / 256), (unsigned)(H - 2)), block(256);
        // With four nodes per thread the first iteration also writes an 8-bit
        // copy of the costs (and a flag for costs without an exact 8-bit
        // form), and the later ones read the copy where it holds every cost.
        uint8_t *d8 = nullptr;
        int *flag = nullptr;
        const bool c8 = x4 && iters > 1;
        if (c8 && e == hipSuccess) {
            if ((e = scratch_alloc((void **)&d8, (size_t)W * H * L, st)) == hipSuccess &&
                (e = scratch_alloc((void **)&flag, sizeof(int), st)) == hipSuccess)
                e = ginf::launch_fill_u32(flag, 1, 0u, st);
        }
        for (int t = 0; t < iters && e == hipSuccess; ++t) {
            if (x4) {
                if (!c8)
                    hipLaunchKernelGGL((k_icm_sync_lds_x4<T, 0>), dim3(lds_grid), dim3(bs), lds_bytes_x4, st, data, d8,
                                       flag, table, W, H, L, bufs[t & 1], bufs[(t + 1) & 1]);
                else if (t == 0)
                    hipLaunchKernelGGL((k_icm_sync_lds_x4<T, 1>), dim3(lds_grid), dim3(bs), lds_bytes_x4, st, data, d8,
                                       flag, table, W, H, L, bufs[t & 1], bufs[(t + 1) & 1]);
                else
                    hipLaunchKernelGGL((k_icm_sync_lds_x4<T, 2>), dim3(lds_grid), dim3(bs), lds_bytes_x4, st, data, d8,
                                       flag, table, W, H, L, bufs[t & 1], bufs[(t + 1) & 1]);
            } else if (lds)
                hipLaunchKernelGGL(k_icm_sync_lds<T>, dim3(lds_grid), dim3(1024), lds_bytes, st, data, table, W, H, L,
                                   bufs[t & 1], bufs[(t + 1) & 1]);
            else
                hipLaunchKernelGGL(k_icm_sync<T>, grid, block, 0, st, data, table, W, H, L, bufs[t & 1],
                                   bufs[(t + 1) & 1]);
            e = hipGetLastError();
        }
        hipError_t f = hipFreeAsync(tmp, st);
        if (d8 && f == hipSuccess) f = hipFreeAsync(d8, st);
        if (flag && f == hipSuccess) f = hipFreeAsync(flag, st);
        return e != hipSuccess ? e : f;
    }
    const int nd = (W - 2) + (H - 2) - 1;  // anti-diagonals of the interior
    const int longest = min(W - 2, H - 2);
    const int steps = nd + 2 * (iters - 1);
    // Node-major copy of the costs (one read and one write of the volume).
    const int64_t P = (int64_t)W * H;
    T *nm = nullptr;
    if ((e = scratch_alloc((void **)&nm, (size_t)P * L * sizeof(T), st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_costs_node_major<T>, dim3((unsigned)((P + 63) / 64), (unsigned)((L + 31) / 32)), dim3(256),
                       0, st, data, P, L, nm);
    e = hipGetLastError();
    const size_t nm_bytes = (size_t)P * L * sizeof(T);
    if (lds && nm_bytes < ((size_t)1 << 32)) {
        // tiles: one launch per tile anti-diagonal and pipelined iteration
        constexpr int TS = GINF_ICM_TILE;
        const size_t tile_lds = (size_t)L * kLdsMaxLabels * sizeof(T);
        const int ntx = (W - 2 + TS - 1) / TS, nty = (H - 2 + TS - 1) / TS, ndt = ntx + nty - 1;
        const int tsteps = ndt + 2 * (iters - 1);
        const dim3 block(TS * kTileLanes);
        for (int s = 0; s < tsteps && e == hipSuccess; ++s) {
            const int t_lo = s < ndt ? 0 : (s - ndt) / 2 + 1;
            const int t_hi = min(iters - 1, s / 2);
            if (t_lo > t_hi) continue;
            const dim3 grid((unsigned)min(ntx, nty), (unsigned)(t_hi - t_lo + 1));
            hipLaunchKernelGGL((k_icm_async_tiles<T, TS, kTileLanes>), grid, block, tile_lds, st, nm, table, W, H, L, result, s,
                               t_lo, ntx, nty, (uint32_t)nm_bytes);
            e = hipGetLastError();
        }
        const hipError_t f = hipFreeAsync(nm, st);
        return e != hipSuccess ? e : f;
    }
    for (int s = 0; s < steps && e == hipSuccess; ++s) {
        // iterations with 0 <= s - 2t < nd
        const int t_lo = s < nd ? 0 : (s - nd) / 2 + 1;
        const int t_hi = min(iters - 1, s / 2);
        if (t_lo > t_hi) continue;
        // (the table stays in global memory here: a launch has too few nodes
        // to repay staging it in every block; measured 2x slower with LDS)
        const int64_t lanes = (int64_t)longest * kAsyncLanes;
        const dim3 grid((unsigned)((lanes + 255) / 256), (unsigned)(t_hi - t_lo + 1)), block(256);
        hipLaunchKernelGGL(k_icm_async_lanes<T>, grid, block, 0, st, nm, table, W, H, L, result, s, t_lo, nd);
        e = hipGetLastError();
    }
    const hipError_t f = hipFreeAsync(nm, st);
    return e != hipSuccess ? e : f;
}

}  // namespace

extern "C" {

int ginf_icm(int width, int height, int num_labels, int dtype, const void *data_dev, const void *table_dev,
             const int32_t *initial_dev, int iters, int mode, int32_t *result_dev, void *stream) {
    using ginf::set_last_error;
    if (!data_dev || !table_dev || !initial_dev || !result_dev) return set_last_error(GINF_ERR_INVALID, "NULL argument");
    if (width < 1 || height < 1 || num_labels < 1 || iters < 0)
        return set_last_error(GINF_ERR_INVALID, "bad sizes or iteration count");
    if (mode != GINF_ICM_SYNC && mode != GINF_ICM_ASYNC) return set_last_error(GINF_ERR_INVALID, "bad ICM mode");
    if ((int64_t)width * height > ((int64_t)1 << 30) || height - 2 > 65535)
        return set_last_error(GINF_ERR_UNSUPPORTED, "grid too large");
    if (mode == GINF_ICM_ASYNC && iters > 65535)
        return set_last_error(GINF_ERR_UNSUPPORTED, "more than 65535 asynchronous iterations");
    hipError_t e;
    if (dtype == GINF_F32)
        e = icm_t((const float *)data_dev, (const float *)table_dev, width, height, num_labels, initial_dev, iters,
                  mode, result_dev, (hipStream_t)stream);
    else if (dtype == GINF_I32)
        e = icm_t((const int *)data_dev, (const int *)table_dev, width, height, num_labels, initial_dev, iters, mode,
                  result_dev, (hipStream_t)stream);
    else
        return set_last_error(GINF_ERR_INVALID, "dtype must be GINF_F32 or GINF_I32");
    return e == hipSuccess ? GINF_OK : set_last_error(GINF_ERR_HIP, hipGetErrorString(e));
}

int ginf_max_prior(int width, int height, int num_labels, int dtype, const void *data_dev, int32_t *result_dev,
                   void *stream) {
    if (!data_dev || !result_dev || width < 1 || height < 1 || num_labels < 1) return GINF_ERR_INVALID;
    const int64_t n = (int64_t)width * height;
    if (n % 4 == 0 && ((uintptr_t)data_dev & 15) == 0 && ((uintptr_t)result_dev & 15) == 0 &&
        (dtype == GINF_F32 || dtype == GINF_I32)) {
        const unsigned qb = (unsigned)std::min<int64_t>((n / 4 + 255) / 256, 65536);
        if (dtype == GINF_F32)
            hipLaunchKernelGGL(k_max_prior_x4<float>, dim3(qb), dim3(256), 0, (hipStream_t)stream,
                               (const float *)data_dev, n, num_labels, result_dev);
        else
            hipLaunchKernelGGL(k_max_prior_x4<int>, dim3(qb), dim3(256), 0, (hipStream_t)stream, (const int *)data_dev,
                               n, num_labels, result_dev);
        return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
    }
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    if (dtype == GINF_F32)
        hipLaunchKernelGGL(k_max_prior<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const float *)data_dev,
                           width, height, num_labels, result_dev);
    else if (dtype == GINF_I32)
        hipLaunchKernelGGL(k_max_prior<int>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const int *)data_dev,
                           width, height, num_labels, result_dev);
    else
        return GINF_ERR_INVALID;
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}

// Host-pointer entries for the C++ templates (decodeIcmSync / gpuDecodeIcm /
// decodeIcmAsync, decodeMaxPrior): copy in, run, copy the labels back.
int ginf_icm_host(int width, int height, int num_labels, int dtype, const void *data_host, const void *table_host,
                  const int32_t *initial_host, int iters, int mode, int32_t *result_host) {
    if (!data_host || !table_host || !initial_host || !result_host) return GINF_ERR_INVALID;
    if (width < 1 || height < 1 || num_labels < 1) return GINF_ERR_INVALID;
    const size_t nd = (size_t)width * height * num_labels * 4, nt = (size_t)num_labels * num_labels * 4;
    const size_t nl = (size_t)width * height * sizeof(int32_t);
    void *dd = nullptr, *dt = nullptr, *di = nullptr, *dr = nullptr;
    hipError_t e = hipMalloc(&dd, nd);
    if (e == hipSuccess) e = hipMalloc(&dt, nt);
    if (e == hipSuccess) e = hipMalloc(&di, nl);
    if (e == hipSuccess) e = hipMalloc(&dr, nl);
    if (e == hipSuccess) e = hipMemcpy(dd, data_host, nd, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, table_host, nt, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(di, initial_host, nl, hipMemcpyHostToDevice);
    int rc = GINF_ERR_HIP;
    if (e == hipSuccess) {
        rc = ginf_icm(width, height, num_labels, dtype, dd, dt, (const int32_t *)di, iters, mode, (int32_t *)dr, nullptr);
        if (rc == GINF_OK) e = hipMemcpy(result_host, dr, nl, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = GINF_ERR_HIP;
    }
    for (void *p : {dd, dt, di, dr})
        if (p) (void)hipFree(p);
    return rc;
}

int ginf_max_prior_host(int width, int height, int num_labels, int dtype, const void *data_host, int32_t *result_host) {
    if (!data_host || !result_host || width < 1 || height < 1 || num_labels < 1) return GINF_ERR_INVALID;
    const size_t nd = (size_t)width * height * num_labels * 4, nl = (size_t)width * height * sizeof(int32_t);
    void *dd = nullptr, *dr = nullptr;
    hipError_t e = hipMalloc(&dd, nd);
    if (e == hipSuccess) e = hipMalloc(&dr, nl);
    if (e == hipSuccess) e = hipMemcpy(dd, data_host, nd, hipMemcpyHostToDevice);
    int rc = GINF_ERR_HIP;
    if (e == hipSuccess) {
        rc = ginf_max_prior(width, height, num_labels, dtype, dd, (int32_t *)dr, nullptr);
        if (rc == GINF_OK) e = hipMemcpy(result_host, dr, nl, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = GINF_ERR_HIP;
    }
    if (dd) (void)hipFree(dd);
    if (dr) (void)hipFree(dr);
    return rc;
}

}  // extern "C"
