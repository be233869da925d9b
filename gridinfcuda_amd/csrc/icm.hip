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
// Costs stay in the reference layout [l][y][x]: a wave reads 64 adjacent x
// of one label plane per load.  The L x L table V(a,b) at a + L*b is read
// through the cache (column f of V is contiguous in a).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "bp_kernels.h"
#include "ginf_bp.h"

namespace {

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
    T best = T(0);
    int arg = 0;
    for (int l = 0; l < L; ++l) {
        T c = data[(int64_t)l * plane + node];
        c = c + vS[l];
        c = c + vE[l];
        c = c + vN[l];
        c = c + vW[l];
        if (l == 0 || c < best) { best = c; arg = l; }
    }
    return arg;
}

template <typename T>
__global__ void k_icm_sync(const T *__restrict__ data, const T *__restrict__ table, int W, int H, int L,
                           const int32_t *__restrict__ old, int32_t *__restrict__ out) {
    const int x = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = 1 + blockIdx.y;
    if (x > W - 2 || y > H - 2) return;
    const int64_t n = (int64_t)y * W + x;
    out[n] = icm_best(data, table, (int64_t)W * H, n, L, old[n + W], old[n + 1], old[n - W], old[n - 1]);
}

// Launch s: iteration t = t0 + blockIdx.y updates anti-diagonal j = s - 2t
// (x + y = 2 + j) of the interior, in place.
template <typename T>
__global__ void k_icm_async_step(const T *__restrict__ data, const T *__restrict__ table, int W, int H, int L,
                                 int32_t *__restrict__ lab, int s, int t0, int nd) {
    const int t = t0 + blockIdx.y;
    const int j = s - 2 * t;
    if (j < 0 || j >= nd) return;
    const int k = 2 + j;
    const int xs = max(1, k - (H - 2)), xe = min(W - 2, k - 1);
    const int x = xs + blockIdx.x * blockDim.x + threadIdx.x;
    if (x > xe) return;
    const int y = k - x;
    const int64_t n = (int64_t)y * W + x;
    lab[n] = icm_best(data, table, (int64_t)W * H, n, L, lab[n + W], lab[n + 1], lab[n - W], lab[n - 1]);
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

template <typename T>
hipError_t icm_t(const T *data, const T *table, int W, int H, int L, const int32_t *initial, int iters, int mode,
                 int32_t *result, hipStream_t st) {
    const size_t nl = (size_t)W * H * sizeof(int32_t);
    hipError_t e = hipSuccess;
    if (result != initial) e = hipMemcpyAsync(result, initial, nl, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess || iters == 0 || W < 3 || H < 3) return e;
    if (mode == GINF_ICM_SYNC) {
        int32_t *tmp = nullptr;
        if ((e = hipMallocAsync((void **)&tmp, nl, st)) != hipSuccess) return e;
        e = hipMemcpyAsync(tmp, initial, nl, hipMemcpyDeviceToDevice, st);
        // The last iteration writes `result`.
        int32_t *bufs[2] = {(iters & 1) ? tmp : result, (iters & 1) ? result : tmp};
        const dim3 grid((unsigned)((W - 2 + 255) / 256), (unsigned)(H - 2)), block(256);
        for (int t = 0; t < iters && e == hipSuccess; ++t) {
            hipLaunchKernelGGL(k_icm_sync<T>, grid, block, 0, st, data, table, W, H, L, bufs[t & 1], bufs[(t + 1) & 1]);
            e = hipGetLastError();
        }
        const hipError_t f = hipFreeAsync(tmp, st);
        return e != hipSuccess ? e : f;
    }
    const int nd = (W - 2) + (H - 2) - 1;  // anti-diagonals of the interior
    const int longest = min(W - 2, H - 2);
    const int steps = nd + 2 * (iters - 1);
    for (int s = 0; s < steps && e == hipSuccess; ++s) {
        // iterations with 0 <= s - 2t < nd
        const int t_lo = s < nd ? 0 : (s - nd) / 2 + 1;
        const int t_hi = min(iters - 1, s / 2);
        if (t_lo > t_hi) continue;
        const dim3 grid((unsigned)((longest + 255) / 256), (unsigned)(t_hi - t_lo + 1)), block(256);
        hipLaunchKernelGGL(k_icm_async_step<T>, grid, block, 0, st, data, table, W, H, L, result, s, t_lo, nd);
        e = hipGetLastError();
    }
    return e;
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
