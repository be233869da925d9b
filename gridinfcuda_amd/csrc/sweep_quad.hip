// Quad-per-node O(L) half-sweep kernels (exact reference operation order).
#include "sweep_common.h"

namespace ginf {
namespace {

// Broadcast lane k of each group of four lanes (DPP quad_perm [k,k,k,k]).
template <int K>
__device__ __forceinline__ float quad_bcast(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), K * 0x55, 0xF, 0xF, false));
}
template <int K>
__device__ __forceinline__ int quad_bcast(int v) {
    return __builtin_amdgcn_update_dpp(0, v, K * 0x55, 0xF, 0xF, false);
}

// Red-black half-sweep for the O(L) models with four lanes per node, lane d
// owning direction d: every lane forms the belief in the reference order
// (D + in_S + in_E + in_N + in_W, the four incoming messages shared across the
// quad by DPP broadcasts) and then runs bpSendMsg for its own direction with
// the reference's sequential scans and left-fold normaliser, so every
// operation -- FP32 mean mode included -- rounds exactly as src/bp.cc:26-54.
//   LMAX  register capacity in labels (L <= LMAX, checked at run time)
// Layout: one lane per node (G = 1 planes).
template <typename T, int LMAX, bool ZEROIN, bool LABELS>
__global__ __launch_bounds__(256) void k_sweep_quad(const SweepArgs a) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = gid >> 2;
    const int d = gid & 3;
    if (slot >= a.nslots) return;
    const int y = slot / a.whp;
    const int x = 2 * (slot - y * a.whp) + ((y + a.color) & 1);
    if (y < 1 || y > a.h - 2 || x < 1 || x > a.w - 2) return;  // whole quads leave together
    const int L = a.num_labels;
    const T s = Num<T>::s(a);
    const T tau = Num<T>::tau(a);
    const int64_t stride = a.stride;
    const T *dp = (const T *)a.data + slot;
    const T *ip = nullptr;
    if (!ZEROIN) {
        int nx = x + kDX[d], ny = y + kDY[d];
        if (a.src == kSrcCoarse) { nx >>= 1; ny >>= 1; }
        const int od = (d + 2) & 3;
        const int c = (nx + ny) & 1;
        ip = (const T *)a.in[c][od] + ((int64_t)ny * a.in_whp + (nx >> 1));
    }

    T f[LMAX];
    T mu = T(0), best = T(0);
    int arg = 0;
#pragma unroll
    for (int i = 0; i < LMAX; ++i) {
        if (i < L) {
            const T m = ZEROIN ? T(0) : ip[i * a.in_stride];
            T b = dp[i * stride];
            b = b + quad_bcast<0>(m);
            b = b + quad_bcast<1>(m);
            b = b + quad_bcast<2>(m);
            b = b + quad_bcast<3>(m);
            if (LABELS) {
                if (i == 0) { best = b; arg = 0; }
                else if (b < best) { best = b; arg = i; }
            }
            const T h = b - m;
            if (i == 0) { f[0] = h; mu = h; }
            else { f[i] = tmin(h, f[i - 1] + s); mu = tmin(mu, h); }
        }
    }
#pragma unroll
    for (int i = LMAX - 2; i >= 0; --i)
        if (i <= L - 2) f[i] = tmin(f[i], f[i + 1] + s);
    const T cap = mu + tau;
    T acc = T(0), mn = T(0);
#pragma unroll
    for (int i = 0; i < LMAX; ++i) {
        if (i < L) {
            f[i] = tmin(f[i], cap);
            acc = acc + f[i];
            mn = (i == 0) ? f[0] : tmin(mn, f[i]);
        }
    }
    const T nu = (a.norm == 0) ? acc / T(L) : mn;
    T *op = (T *)a.out[d] + slot;
#pragma unroll
    for (int i = 0; i < LMAX; ++i)
        if (i < L) op[i * stride] = f[i] - nu;
    if (LABELS && d == 0) a.labels[(int64_t)y * a.label_w + x] = arg;
}

template <typename T, int LMAX>
hipError_t launch_quad_l(const SweepArgs &a, bool zeroin, bool labels, hipStream_t st) {
    const int64_t threads = (int64_t)a.nslots * 4;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    if (zeroin) { if (labels) hipLaunchKernelGGL((k_sweep_quad<T, LMAX, true, true>), grid, block, 0, st, a);
                  else hipLaunchKernelGGL((k_sweep_quad<T, LMAX, true, false>), grid, block, 0, st, a); }
    else { if (labels) hipLaunchKernelGGL((k_sweep_quad<T, LMAX, false, true>), grid, block, 0, st, a);
           else hipLaunchKernelGGL((k_sweep_quad<T, LMAX, false, false>), grid, block, 0, st, a); }
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
