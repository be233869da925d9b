// HBM calibration probe (measurement hook, no reference counterpart): a
// grid-stride kernel that reads R float4 streams and writes W float4 streams
// (non-temporal stores), the shape of the BP half-sweep's traffic (data + 4
// incoming outbox streams read, 4 outbox streams written: R = 5, W = 4) and
// of a plain copy (R = 1, W = 1).  Its achieved bandwidth is the practical
// ceiling the sweep's roofline fraction is read against.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ginf_bp.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

template <int R, int W>
__global__ __launch_bounds__(256) void k_stream(const v4f *__restrict__ in, v4f *__restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        v4f acc = v4f(0.f);
#pragma unroll
        for (int r = 0; r < R; ++r) acc += in[r * n + i];
#pragma unroll
        for (int w = 0; w < W; ++w) __builtin_nontemporal_store(acc + (float)w, &out[w * n + i]);
    }
}

}  // namespace

extern "C" int ginf_stream_probe(int reads, int writes, const void *in_dev, void *out_dev, int64_t n_vec4, int blocks,
                                 void *stream) {
    if (!in_dev || !out_dev || n_vec4 < 1 || blocks < 1) return GINF_ERR_INVALID;
    const dim3 g((unsigned)blocks), b(256);
    hipStream_t st = (hipStream_t)stream;
    const v4f *in = (const v4f *)in_dev;
    v4f *out = (v4f *)out_dev;
    if (reads == 1 && writes == 1) hipLaunchKernelGGL((k_stream<1, 1>), g, b, 0, st, in, out, n_vec4);
    else if (reads == 5 && writes == 4) hipLaunchKernelGGL((k_stream<5, 4>), g, b, 0, st, in, out, n_vec4);
    else return GINF_ERR_UNSUPPORTED;
    return hipGetLastError() == hipSuccess ? GINF_OK : GINF_ERR_HIP;
}
