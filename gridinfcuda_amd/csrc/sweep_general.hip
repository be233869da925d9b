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
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_sweep_general(int dtype, const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                                hipStream_t st) {
    return dtype == 0 ? launch_general_t<float>(sh, a, zeroin, labels, st)
                      : launch_general_t<int>(sh, a, zeroin, labels, st);
}

}  // namespace ginf
