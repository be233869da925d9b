// Batched persistent half-sweep: one launch moves one colour of every
// instance of a batch.  The grid is sized to the machine (blocks per CU
// chosen by the tuner) and each block walks the work list with a grid
// stride; a work item is a run of slots of one instance, processed 64 slots
// (256 lanes, one lane per node and direction) at a time with the quad body,
// dispatched on the instance's label capacity.  Instances keep the
// single-grid layout, so results equal the single-grid kernel bit for bit.
#include "quad_body.h"

namespace ginf {
namespace {

template <typename T, bool ZEROIN, bool LABELS, bool FULL>
__device__ __forceinline__ void batch_node(const SweepArgs &a, int slot, int d) {
    switch (a.lmax) {
        case 16: quad_node<T, 16, ZEROIN, LABELS, FULL>(a, slot, d); break;
        case 32: quad_node<T, 32, ZEROIN, LABELS, FULL>(a, slot, d); break;
        case 64: quad_node<T, 64, ZEROIN, LABELS, FULL>(a, slot, d); break;
        case 96: quad_node<T, 96, ZEROIN, LABELS, FULL>(a, slot, d); break;
        case 128: quad_node<T, 128, ZEROIN, LABELS, FULL>(a, slot, d); break;
        default: break;
    }
}

template <typename T, bool ZEROIN, bool LABELS>
__global__ __launch_bounds__(256) void k_sweep_batch(const SweepArgs *__restrict__ args,
                                                     const BatchItem *__restrict__ items, int nitems) {
    const int d = threadIdx.x & 3;
    const int lane_slot = threadIdx.x >> 2;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BatchItem w = items[it];
        const SweepArgs a = args[w.inst];
        const bool full = a.num_labels == a.lmax;
        for (int s0 = 0; s0 < w.nslots; s0 += 64) {
            if (lane_slot >= w.nslots - s0) break;
            const int slot = w.slot0 + s0 + lane_slot;
            if (full) batch_node<T, ZEROIN, LABELS, true>(a, slot, d);
            else batch_node<T, ZEROIN, LABELS, false>(a, slot, d);
        }
    }
}

}  // namespace

hipError_t launch_sweep_batch(int dtype, const SweepArgs *args_dev, const BatchItem *items_dev, int nitems, int grid,
                              bool zeroin, bool labels, hipStream_t st) {
    if (nitems <= 0) return hipSuccess;
    const dim3 g((unsigned)(grid < nitems ? grid : nitems)), b(256);
#define GINF_BK(TT, Z, LB) hipLaunchKernelGGL((k_sweep_batch<TT, Z, LB>), g, b, 0, st, args_dev, items_dev, nitems)
    if (dtype == 0) {
        if (zeroin) { if (labels) GINF_BK(float, true, true); else GINF_BK(float, true, false); }
        else { if (labels) GINF_BK(float, false, true); else GINF_BK(float, false, false); }
    } else {
        if (zeroin) { if (labels) GINF_BK(int, true, true); else GINF_BK(int, true, false); }
        else { if (labels) GINF_BK(int, false, true); else GINF_BK(int, false, false); }
    }
#undef GINF_BK
    return hipGetLastError();
}

}  // namespace ginf
