// First half-sweep of a seeded level, its incoming messages read straight
// from the network's prediction (quad layout, FP32 messages).  Seeding the
// level's outboxes first (k_inject) writes all eight outbox arrays only for
// this half-sweep to read colour B's four of them back; reading the
// prediction here instead (1/16 of the level's size at 1/4 resolution, so it
// stays in cache) saves that write and that read.  Every incoming value is the
// one k_inject would have stored, so the result is the same bit for bit.
#include "quad_body.h"

namespace ginf {
namespace {

template <int LMAX, bool LABELS, bool FULL, bool D16>
__global__ __launch_bounds__(256) void k_sweep_seeded(const SweepArgs a) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    quad_node<float, LMAX, false, LABELS, FULL, D16, true, D16>(a, gid >> 2, gid & 3);
}

template <int LMAX>
hipError_t launch_seeded_l(const SweepArgs &a, bool labels, bool d16, hipStream_t st) {
    const int64_t threads = (int64_t)a.nslots * 4;
    const int bs = threads < 100000 ? 64 : 256;  // as the quad half-sweep: small levels in 64-lane blocks
    const dim3 grid((unsigned)((threads + bs - 1) / bs)), block(bs);
    const bool full = a.num_labels == LMAX;
#define GINF_SK(LB, F, D) hipLaunchKernelGGL((k_sweep_seeded<LMAX, LB, F, D>), grid, block, 0, st, a)
#define GINF_SK2(LB, F) { if (d16) GINF_SK(LB, F, true); else GINF_SK(LB, F, false); }
    if (full) { if (labels) GINF_SK2(true, true) else GINF_SK2(false, true) }
    else { if (labels) GINF_SK2(true, false) else GINF_SK2(false, false) }
#undef GINF_SK2
#undef GINF_SK
    return hipGetLastError();
}

}  // namespace

hipError_t launch_sweep_seeded(const SweepShape &sh, const SweepArgs &a, bool labels, bool d16, hipStream_t st) {
    if (!sh.quad || a.num_labels % 4) return hipErrorInvalidValue;
    switch (sh.C) {
#define GINF_QL(LM) case LM: return launch_seeded_l<LM>(a, labels, d16, st);
        GINF_QUAD_LIST(GINF_QL)
#undef GINF_QL
        default: return hipErrorInvalidValue;
    }
}

}  // namespace ginf
