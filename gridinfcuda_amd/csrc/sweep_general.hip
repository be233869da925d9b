// O(L^2) half-sweep kernels for the EXPLICIT and TRUNC_QUAD models.
#include "sweep_common.h"

namespace ginf {
namespace {

// Half-sweep for the general O(L^2) models (EXPLICIT, TRUNC_QUAD), one lane
// per node: M(p,d,i) = MIN_j (V(i,j) + b_j) for every d (src/bp.cc:55-62).
template <typename T, int C, bool ZEROIN, bool LABELS>
__global__ __launch_bounds__(256) void k_sweep_general(const SweepArgs a) {
    int slot, q, x, y;
    if (!locate(a, 1, slot, q, x, y)) return;
    const int L = a.num_labels;
    const int64_t stride = a.stride;
    const int64_t self = tile_slot(a, slot, 1);
    const T *dp = (const T *)a.data + self;
    const T *ip[4];
    if (!ZEROIN) incoming_ptrs<T, 1>(a, x, y, 0, ip);
    const T *V = (const T *)a.table;

    T b[C];
    T best = T(0);
    int arg = 0;
#pragma unroll
    for (int i = 0; i < C; ++i) {
        if (i < L) {
            T v = dp[i * stride];
#pragma unroll
            for (int d = 0; d < 4; ++d) v = v + (ZEROIN ? T(0) : ip[d][i * a.in_stride]);
            b[i] = v;
            if (LABELS) {
                if (i == 0) { best = v; arg = 0; }
                else if (v < best) { best = v; arg = i; }
            }
        } else {
            b[i] = T(0);
        }
    }
    for (int i = 0; i < L; ++i) {
        T v = V[i] + b[0];
#pragma unroll
        for (int j = 0; j < C; ++j)
            if (j < L) v = tmin(v, V[i + L * j] + b[j]);
#pragma unroll
        for (int d = 0; d < 4; ++d) ((T *)a.out[d])[self + i * stride] = v;
    }
    if (LABELS) a.labels[(int64_t)y * a.label_w + x] = arg;
}

// The same update for label counts past what one lane's registers hold
// (128 < L <= kWaveLabels): one wave per node.  The belief b is formed
// label by label across the lanes (lane k takes labels k, k + 64, ...) into
// LDS, then lane k computes M(i) for its labels i with the same scan over j
// from V(i,0) + b_0 upwards.  The minimum of a fixed set of values does not
// depend on the order it is taken in, so the messages are exactly those of
// the one-lane kernel and of src/bp.cc:55-62; the label is the lowest index
// of minimal belief, combined across lanes.
constexpr int kWaveLabels = 1024;
constexpr int kWavesPerBlock = 4;

template <typename T, bool ZEROIN, bool LABELS>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_sweep_general_wave(const SweepArgs a) {
    __shared__ T bsh[kWavesPerBlock][kWaveLabels];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = blockIdx.x * kWavesPerBlock + wave;
    if (slot >= a.nslots) return;  // whole waves leave together
    const int y = slot / a.whp;
    const int x = 2 * (slot - y * a.whp) + ((y + a.color) & 1);
    if (y < 1 || y > a.h - 2 || x < 1 || x > a.w - 2) return;
    const int L = a.num_labels;
    const int64_t stride = a.stride;
    const int64_t self = tile_slot(a, slot, 1);
    const T *dp = (const T *)a.data + self;
    const T *ip[4];
    if (!ZEROIN) incoming_ptrs<T, 1>(a, x, y, 0, ip);
    const T *V = (const T *)a.table;
    T *b = bsh[wave];
    T best = T(0);
    int arg = -1;
    for (int i = lane; i < L; i += 64) {
        T v = dp[i * stride];
#pragma unroll
        for (int d = 0; d < 4; ++d) v = v + (ZEROIN ? T(0) : ip[d][i * a.in_stride]);
        b[i] = v;
        if (LABELS && (arg < 0 || v < best)) { best = v; arg = i; }
    }
    // the wave's own LDS writes, read back by other lanes of the wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int i = lane; i < L; i += 64) {
        T v = V[i] + b[0];
        for (int j = 1; j < L; ++j) v = tmin(v, V[i + L * j] + b[j]);
#pragma unroll
        for (int d = 0; d < 4; ++d) ((T *)a.out[d])[self + i * stride] = v;
    }
    if (LABELS) {
        // smaller belief wins, ties go to the lower label
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const T ob = __shfl_xor(best, m, 64);
            const int oa = __shfl_xor(arg, m, 64);
            const bool take = oa >= 0 && (arg < 0 || ob < best || (!(best < ob) && oa < arg));
            if (take) { best = ob; arg = oa; }
        }
        if (lane == 0) a.labels[(int64_t)y * a.label_w + x] = arg;
    }
}

template <typename T>
hipError_t launch_general_wave(const SweepArgs &a, bool zeroin, bool labels, hipStream_t st) {
    if (a.num_labels > kWaveLabels) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((a.nslots + kWavesPerBlock - 1) / kWavesPerBlock)), block(64 * kWavesPerBlock);
    if (zeroin) { if (labels) hipLaunchKernelGGL((k_sweep_general_wave<T, true, true>), grid, block, 0, st, a);
                  else hipLaunchKernelGGL((k_sweep_general_wave<T, true, false>), grid, block, 0, st, a); }
    else { if (labels) hipLaunchKernelGGL((k_sweep_general_wave<T, false, true>), grid, block, 0, st, a);
           else hipLaunchKernelGGL((k_sweep_general_wave<T, false, false>), grid, block, 0, st, a); }
    return hipGetLastError();
}

template <typename T, int C>
hipError_t launch_general_c(const SweepArgs &a, bool zeroin, bool labels, hipStream_t st) {
    const dim3 grid((unsigned)((a.nslots + 255) / 256)), block(256);
    if (zeroin) { if (labels) hipLaunchKernelGGL((k_sweep_general<T, C, true, true>), grid, block, 0, st, a);
                  else hipLaunchKernelGGL((k_sweep_general<T, C, true, false>), grid, block, 0, st, a); }
    else { if (labels) hipLaunchKernelGGL((k_sweep_general<T, C, false, true>), grid, block, 0, st, a);
           else hipLaunchKernelGGL((k_sweep_general<T, C, false, false>), grid, block, 0, st, a); }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_general_t(const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels, hipStream_t st) {
    switch (sh.C) {
        case 16: return launch_general_c<T, 16>(a, zeroin, labels, st);
        case 32: return launch_general_c<T, 32>(a, zeroin, labels, st);
        case 64: return launch_general_c<T, 64>(a, zeroin, labels, st);
        case 128: return launch_general_c<T, 128>(a, zeroin, labels, st);
        default: return sh.C > 128 ? launch_general_wave<T>(a, zeroin, labels, st) : hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_sweep_general(int dtype, const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                                hipStream_t st) {
    return dtype == 0 ? launch_general_t<float>(sh, a, zeroin, labels, st)
                      : launch_general_t<int>(sh, a, zeroin, labels, st);
}

}  // namespace ginf
