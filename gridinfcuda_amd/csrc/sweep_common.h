// Device helpers shared by the half-sweep kernels (one copy per translation
// unit): the collect + O(L) truncated-linear / Potts distance transform +
// normalisation, or the O(L^2) general model.  Layout: see bp_layout.h.
//
// Semantics follow the reference CPU BP (src/bp.cc): collect order S,E,N,W
// (:7-18), h = b - reverse message, forward then backward min-plus scans
// (:36-40), clamp at min(h) + V(0,L-1) (:43-48), mean or min normalisation
// (:50-54); the general model (:55-62) skips the reverse message and the
// normalisation.  MIN below is the reference's (a < b) ? a : b
// (src/include/utils.h:16).  With one lane per node every operation runs in
// the reference's order; with several lanes the scans and sums are split
// into per-lane chunks joined by cross-lane carries, which is exact for
// integer arithmetic and within rounding for non-integer FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>

#include "bp_kernels.h"

namespace ginf {
namespace {

constexpr int kDX[4] = {0, 1, 0, -1};
constexpr int kDY[4] = {1, 0, -1, 0};

// MIN of the reference, (a < b) ? a : b.  For float the hardware minimum
// (v_min_f32) is used: it returns the same value for every pair of operands
// except a NaN operand or a +0/-0 pair, which the reference inputs (finite
// costs and tables without negative zeros) never produce, because no
// operation of the algorithm yields -0 or NaN from such inputs.  Build with
// -DGINF_STRICT_MIN to get the literal compare-and-select instead.
template <typename T>
__device__ __forceinline__ T tmin(T a, T b) { return (a < b) ? a : b; }
#ifndef GINF_STRICT_MIN
template <>
__device__ __forceinline__ float tmin<float>(float a, float b) { return __builtin_fminf(a, b); }
#endif

template <typename T> struct Num;
template <> struct Num<float> {
    __device__ static __forceinline__ float top() { return __builtin_huge_valf(); }
    __device__ static __forceinline__ float s(const SweepArgs &a) { return a.s_f; }
    __device__ static __forceinline__ float tau(const SweepArgs &a) { return a.tau_f; }
};
template <> struct Num<int> {
    __device__ static __forceinline__ int top() { return INT_MAX; }
    __device__ static __forceinline__ int s(const SweepArgs &a) { return a.s_i; }
    __device__ static __forceinline__ int tau(const SweepArgs &a) { return a.tau_i; }
};

template <int G, typename T>
__device__ __forceinline__ T lane_up(T v, int off) { return __shfl_up(v, off, G); }
template <int G, typename T>
__device__ __forceinline__ T lane_down(T v, int off) { return __shfl_down(v, off, G); }
template <int G, typename T>
__device__ __forceinline__ T lane_xor(T v, int m) { return __shfl_xor(v, m, G); }

// Element of (slot n, plane 0, lane 0) of a state array of the plan
// (slot_elem of bp_layout.h with the tile size taken from the arguments).
__device__ __forceinline__ int64_t tile_slot(const SweepArgs &a, int64_t n, int G) {
    return (n >> kTileShift) * a.tile + (n & (kTileSlots - 1)) * G;
}

// Locate the node this lane group works on.  Returns false for slots that
// are padding, border nodes or beyond the level (those never move).
__device__ __forceinline__ bool locate(const SweepArgs &a, int G, int &slot, int &q, int &x, int &y) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    slot = gid / G;
    q = gid - slot * G;
    if (slot >= a.nslots) return false;
    y = slot / a.whp;
    const int k = slot - y * a.whp;
    x = 2 * k + ((y + a.color) & 1);
    return y >= 1 && y <= a.h - 2 && x >= 1 && x <= a.w - 2;
}

// Pointers to register plane 0 of the four incoming messages of node (x,y):
// in_d is the outbox of neighbour p + dir(d) in direction OPP(d), read at this
// level (kSrcSame) or at the parent of that neighbour one level up
// (kSrcCoarse, the outbox-frame upsample of src/bp.cc:129-139 fused in).
template <typename T, int G>
__device__ __forceinline__ void incoming_ptrs(const SweepArgs &a, int x, int y, int q, const T *ip[4]) {
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        int nx = x + kDX[d], ny = y + kDY[d];
        if (a.src == kSrcCoarse) { nx >>= 1; ny >>= 1; }
        const int od = (d + 2) & 3;
        const int c = (nx + ny) & 1;
        const T *base = (const T *)(c ? a.in[1][od] : a.in[0][od]);
        ip[d] = base + tile_slot(a, (int64_t)ny * a.in_whp + (nx >> 1), G) + q;
    }
}


}  // namespace
}  // namespace ginf
