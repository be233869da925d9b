// Quad-per-node O(L) half-sweep kernels (exact reference operation order).
#include "quad_body.h"

namespace ginf {
namespace {

template <typename T, int LMAX, bool ZEROIN, bool LABELS, bool FULL>
__global__ __launch_bounds__(256) void k_sweep_quad(const SweepArgs a) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    quad_node<T, LMAX, ZEROIN, LABELS, FULL>(a, gid >> 2, gid & 3);
}

template <typename T, int LMAX>
hipError_t launch_quad_l(const SweepArgs &a, bool zeroin, bool labels, hipStream_t st) {
    const int64_t threads = (int64_t)a.nslots * 4;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    const bool full = a.num_labels == LMAX;
#define GINF_QK(Z, LB, F) hipLaunchKernelGGL((k_sweep_quad<T, LMAX, Z, LB, F>), grid, block, 0, st, a)
    if (full) {
        if (zeroin) { if (labels) GINF_QK(true, true, true); else GINF_QK(true, false, true); }
        else { if (labels) GINF_QK(false, true, true); else GINF_QK(false, false, true); }
    } else {
        if (zeroin) { if (labels) GINF_QK(true, true, false); else GINF_QK(true, false, false); }
        else { if (labels) GINF_QK(false, true, false); else GINF_QK(false, false, false); }
    }
#undef GINF_QK
    return hipGetLastError();
}

template <typename T>
hipError_t launch_quad_t(const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels, hipStream_t st) {
    switch (sh.C) {
#define GINF_QL(LM) case LM: return launch_quad_l<T, LM>(a, zeroin, labels, st);
        GINF_QUAD_LIST(GINF_QL)
#undef GINF_QL
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_sweep_quad(int dtype, const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                             hipStream_t st) {
    return dtype == 0 ? launch_quad_t<float>(sh, a, zeroin, labels, st) : launch_quad_t<int>(sh, a, zeroin, labels, st);
}

}  // namespace ginf
