// Alpha-beta swap on gfx950 (reference decodeAlphaBeta, src/graphcut.cc:194-279,
// int costs only as in the reference).
//
// For every label pair (alpha < beta) the swap graph is built as
// graphCutConstructGraph does (graphcut.cc:9-46) -- n-links V(alpha,beta)
// between neighbours labelled alpha or beta, source t-link D(alpha) + sum
// V(alpha, f_q) and sink t-link D(beta) + sum V(beta, f_q) over the
// neighbours q outside the pair -- but from a fresh state (the reference
// keeps the previous pair's excess, residues and heights at nodes outside the
// pair).  A parallel push-relabel computes a maximum preflow:
//   * global relabel: heights = BFS distance to the sink over residual edges,
//     relaxed in place until nothing changes (unreachable = kInf);
//   * push: an active node (excess, height < kInf) pushes to the sink when its
//     height is 1 and to neighbours one level lower; excess and residues are
//     updated with atomics.  Heights are fixed during a push launch, so two
//     nodes never push along one edge in opposite directions, a node's own
//     residues only grow under it, and its excess only grows, which makes the
//     snapshot it pushes from safe;
//   * relabel: nodes without an admissible edge move to 1 + the lowest
//     residual neighbour (1 with sink residue), read from the old heights and
//     written to a second buffer.
// After the last global relabel the nodes with a finite height are exactly
// those that can reach the sink in the residual graph -- the same set for any
// maximum preflow -- and take alpha; the rest of the pair take beta
// (graphCutLabelFromFlow, graphcut.cc:66-79).  The move is kept when the
// labelling cost (FP64 on the device, exact for int costs) drops
// (graphcut.cc:263-272).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <utility>

#include "bp_kernels.h"
#include "ginf_bp.h"

namespace {

constexpr int kDX[4] = {0, 1, 0, -1};  // S, E, N, W
constexpr int kDY[4] = {1, 0, -1, 0};

struct Flow {
    int *e, *rs, *r, *h, *h2, *in, *counter;  // counter[0]: active nodes, counter[1]: height changes
    int W, H, inf;
};

__global__ void k_ab_build(const int *__restrict__ data, const int *__restrict__ table, int L,
                           const int32_t *__restrict__ f, int alpha, int beta, Flow g) {
    const int64_t N = (int64_t)g.W * g.H;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(n % g.W), y = (int)(n / g.W);
        const int fl = f[n];
        const bool in = fl == alpha || fl == beta;
        int ac = 0, bc = 0;
        for (int d = 0; d < 4; ++d) {
            const int nx = x + kDX[d], ny = y + kDY[d];
            int cap = 0;
            if (in && nx >= 0 && ny >= 0 && nx < g.W && ny < g.H) {
                int fq = f[(int64_t)ny * g.W + nx];
                fq = fq < 0 ? 0 : (fq >= L ? L - 1 : fq);  // checked on the host; keeps reads in the table
                if (fq == alpha || fq == beta) cap = table[alpha + L * beta];
                else { ac += table[alpha + L * fq]; bc += table[beta + L * fq]; }
            }
            g.r[4 * n + d] = cap;
        }
        g.in[n] = in;
        g.e[n] = in ? data[(int64_t)alpha * N + n] + ac : 0;
        g.rs[n] = in ? data[(int64_t)beta * N + n] + bc : 0;
        g.h[n] = (in && g.rs[n] > 0) ? 1 : g.inf;
    }
}

__global__ void k_clear(int *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// In-place relaxation steps of the BFS distances to the sink: GINF_AB_BFS_INNER
// sweeps over the nodes per launch, separated by block barriers.  Heights only
// decrease and every value read is an upper bound of the distance, so any
// interleaving converges to the same fixed point (the exact distances); a
// launch that changes nothing has reached it.
// 384x288x16, one cycle, alternating runs on one box: 1 sweep per launch
// 49.9-51.9 ms, 2 sweeps 43.8 (batches of 4), 4 sweeps 40.8-46.1 (batches of
// 2), 46.4 (batches of 4), 52.1 (batches of 1), 8 sweeps 51.2 (batches of 1).
#ifndef GINF_AB_BFS_INNER
#define GINF_AB_BFS_INNER 4
#endif
__global__ void k_ab_bfs(Flow g) {
    const int64_t N = (int64_t)g.W * g.H;
    int changed = 0;
    for (int it = 0; it < GINF_AB_BFS_INNER; ++it) {
        if (it) __syncthreads();
        for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N;
             n += (int64_t)gridDim.x * blockDim.x) {
            if (!g.in[n]) continue;
            const int x = (int)(n % g.W), y = (int)(n / g.W);
            const int cur = __hip_atomic_load(&g.h[n], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int best = cur;
            for (int d = 0; d < 4; ++d) {
                const int nx = x + kDX[d], ny = y + kDY[d];
                if (nx < 0 || ny < 0 || nx >= g.W || ny >= g.H || g.r[4 * n + d] <= 0) continue;
                const int hv =
                    __hip_atomic_load(&g.h[(int64_t)ny * g.W + nx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (hv < g.inf && hv + 1 < best) best = hv + 1;
            }
            if (best < cur) {
                __hip_atomic_store(&g.h[n], best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                changed = 1;
            }
        }
    }
    if (changed) atomicAdd(&g.counter[1], 1);
}

__global__ void k_ab_push(Flow g) {
    const int64_t N = (int64_t)g.W * g.H;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (int64_t)gridDim.x * blockDim.x) {
        const int hu = g.h[n];
        int ex = __hip_atomic_load(&g.e[n], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!g.in[n] || ex <= 0 || hu >= g.inf) continue;
        if (hu == 1 && g.rs[n] > 0) {
            const int dl = min(ex, g.rs[n]);
            g.rs[n] -= dl;
            atomicSub(&g.e[n], dl);
            ex -= dl;
        }
        const int x = (int)(n % g.W), y = (int)(n / g.W);
        for (int d = 0; d < 4 && ex > 0; ++d) {
            const int nx = x + kDX[d], ny = y + kDY[d];
            if (nx < 0 || ny < 0 || nx >= g.W || ny >= g.H) continue;
            const int64_t v = (int64_t)ny * g.W + nx;
            const int cap = __hip_atomic_load(&g.r[4 * n + d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cap <= 0 || hu != g.h[v] + 1) continue;
            const int dl = min(ex, cap);
            atomicSub(&g.r[4 * n + d], dl);
            atomicAdd(&g.r[4 * v + ((d + 2) & 3)], dl);
            atomicSub(&g.e[n], dl);
            atomicAdd(&g.e[v], dl);
            ex -= dl;
        }
    }
}

// Relabel from the old heights (g.h) into g.h2; counts the nodes still active.
__global__ void k_ab_relabel(Flow g) {
    const int64_t N = (int64_t)g.W * g.H;
    int active = 0;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (int64_t)gridDim.x * blockDim.x) {
        const int hu = g.h[n];
        int nh = hu;
        if (g.in[n] && g.e[n] > 0 && hu < g.inf) {
            const int x = (int)(n % g.W), y = (int)(n / g.W);
            bool admissible = g.rs[n] > 0 && hu == 1;
            int low = g.rs[n] > 0 ? 0 : g.inf;
            for (int d = 0; d < 4; ++d) {
                const int nx = x + kDX[d], ny = y + kDY[d];
                if (nx < 0 || ny < 0 || nx >= g.W || ny >= g.H || g.r[4 * n + d] <= 0) continue;
                const int hv = g.h[(int64_t)ny * g.W + nx];
                if (hu == hv + 1) admissible = true;
                low = min(low, hv);
            }
            if (!admissible) nh = low >= g.inf ? g.inf : low + 1;
            if (nh < g.inf) ++active;
        }
        g.h2[n] = nh;
    }
    if (active) atomicAdd(&g.counter[0], active);
}

__global__ void k_ab_label(const int32_t *__restrict__ f, int32_t *__restrict__ out, Flow g, int alpha, int beta) {
    const int64_t N = (int64_t)g.W * g.H;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (int64_t)gridDim.x * blockDim.x)
        out[n] = g.in[n] ? (g.h[n] < g.inf ? alpha : beta) : f[n];
}

// Keep-or-reject of a pair's move on the device (graphcut.cc:263-272): the
// move is kept when the new labelling cost is lower; *keep tells k_ab_commit
// whether to copy the candidate labels over the current ones.
__global__ void k_ab_decide(const double *__restrict__ en, double *__restrict__ best, int *__restrict__ moves,
                            int *__restrict__ keep) {
    const bool k = *en < *best;
    if (k) { *best = *en; *moves += 1; }
    *keep = k ? 1 : 0;
}

__global__ void k_ab_commit(const int *__restrict__ keep, const int32_t *__restrict__ cand, int32_t *__restrict__ cur,
                            int64_t N) {
    if (*keep == 0) return;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (int64_t)gridDim.x * blockDim.x)
        cur[n] = cand[n];
}

struct Ctx {
    hipStream_t st;
    unsigned blocks;
    int *host;  // pinned readback of the counters
};

hipError_t read_counters(Ctx &c, Flow &g) {
    hipError_t e = hipMemcpyAsync(c.host, g.counter, 2 * sizeof(int), hipMemcpyDeviceToHost, c.st);
    if (e == hipSuccess) e = hipStreamSynchronize(c.st);
    return e;
}

// Launches between host checks of the convergence counters: BFS launches of
// the global relabel, and push/relabel rounds of the maximum preflow.  The
// in-place BFS converges in few launches, so its batch is short (384x288x16,
// one cycle, one sweep per BFS launch and a counter clear before every
// launch: 64.0 ms with batches 4/8, 82.8-87.2 with 8/8, 92.7 with 2/4; with
// four sweeps per BFS launch, batches of 2 BFS launches measured best).
#ifndef GINF_AB_BFS_BATCH
#define GINF_AB_BFS_BATCH 2
#endif
#ifndef GINF_AB_ROUNDS
#define GINF_AB_ROUNDS 8
#endif

// Heights := exact BFS distances to the sink over residual edges.
hipError_t global_relabel(Ctx &c, Flow &g) {
    hipError_t e = hipSuccess;
    // Seeds: distance 1 with sink residue, else kInf (partition nodes only).
    const int64_t N = (int64_t)g.W * g.H;
    // Only the last step of a batch is counted (a step that changes nothing
    // is the fixed point), so the counter is cleared just before it.
    for (int64_t guard = 0; guard <= N + 1; guard += GINF_AB_BFS_BATCH) {
        for (int k = 0; k < GINF_AB_BFS_BATCH; ++k) {
            if (k == GINF_AB_BFS_BATCH - 1) hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, c.st, g.counter, 2);
            hipLaunchKernelGGL(k_ab_bfs, dim3(c.blocks), dim3(256), 0, c.st, g);
        }
        if ((e = hipGetLastError()) != hipSuccess || (e = read_counters(c, g)) != hipSuccess) return e;
        if (c.host[1] == 0) return hipSuccess;
    }
    return hipErrorUnknown;  // unreachable: each step lowers some height
}

__global__ void k_ab_reset_heights(Flow g) {
    const int64_t N = (int64_t)g.W * g.H;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (int64_t)gridDim.x * blockDim.x)
        g.h[n] = (g.in[n] && g.rs[n] > 0) ? 1 : g.inf;
}

hipError_t max_preflow(Ctx &c, Flow &g, int64_t max_rounds) {
    hipError_t e;
    const int kRounds = GINF_AB_ROUNDS;
    for (int64_t round = 0; round < max_rounds; round += kRounds) {
        hipLaunchKernelGGL(k_ab_reset_heights, dim3(c.blocks), dim3(256), 0, c.st, g);
        if ((e = global_relabel(c, g)) != hipSuccess) return e;
        for (int k = 0; k < kRounds; ++k) {
            // the host reads the last round's count of active nodes only
            if (k == kRounds - 1) hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, c.st, g.counter, 2);
            hipLaunchKernelGGL(k_ab_push, dim3(c.blocks), dim3(256), 0, c.st, g);
            hipLaunchKernelGGL(k_ab_relabel, dim3(c.blocks), dim3(256), 0, c.st, g);
            std::swap(g.h, g.h2);
        }
        if ((e = hipGetLastError()) != hipSuccess || (e = read_counters(c, g)) != hipSuccess) return e;
        if (c.host[0] == 0) break;
        if (round + kRounds >= max_rounds) return hipErrorUnknown;
    }
    // Final exact distances: finite = can reach the sink.
    hipLaunchKernelGGL(k_ab_reset_heights, dim3(c.blocks), dim3(256), 0, c.st, g);
    return global_relabel(c, g);
}

}  // namespace

extern "C" {

int ginf_alpha_beta(int width, int height, int num_labels, int dtype, const void *data_dev, const void *table_dev,
                    const int32_t *initial_dev, int iters, int32_t *result_dev, int *moves_out, void *stream) {
    using ginf::set_last_error;
    if (!data_dev || !table_dev || !initial_dev || !result_dev || width < 1 || height < 1 || num_labels < 1 ||
        iters < 0)
        return set_last_error(GINF_ERR_INVALID, "NULL argument, bad sizes or iteration count");
    if (dtype != GINF_I32)  // the reference instantiates int only
        return set_last_error(GINF_ERR_UNSUPPORTED, "alpha-beta swap takes int costs only");
    if ((int64_t)width * height > ((int64_t)1 << 28)) return set_last_error(GINF_ERR_UNSUPPORTED, "grid too large");
    const int64_t N = (int64_t)width * height;
    hipStream_t st = (hipStream_t)stream;
    Ctx c{st, (unsigned)std::min<int64_t>((N + 255) / 256, 4096), nullptr};
    int *pool = nullptr;
    int32_t *cand = nullptr;
    double *ebuf = nullptr, *ehost = nullptr;
    const int eblocks = 256;
    hipError_t e = hipMalloc((void **)&pool, sizeof(int) * (size_t)(9 * N + 2));
    if (e == hipSuccess) e = hipMalloc((void **)&cand, sizeof(int32_t) * (size_t)N);
    // ebuf: eblocks partial sums, the candidate's energy, the best energy, then
    // two ints (kept moves, keep flag of the last pair)
    if (e == hipSuccess) e = hipMalloc((void **)&ebuf, sizeof(double) * (eblocks + 3));
    if (e == hipSuccess) e = hipHostMalloc((void **)&c.host, 2 * sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&ehost, sizeof(double) + 2 * sizeof(int), hipHostMallocDefault);
    Flow g{};
    g.W = width;
    g.H = height;
    g.inf = (int)std::min<int64_t>(N + 2, INT32_MAX / 2);
    if (pool) {
        g.e = pool;
        g.rs = pool + N;
        g.h = pool + 2 * N;
        g.h2 = pool + 3 * N;
        g.in = pool + 4 * N;
        g.r = pool + 5 * N;
        g.counter = pool + 9 * N;
    }
    double *en_dev = ebuf ? ebuf + eblocks : nullptr, *best_dev = ebuf ? ebuf + eblocks + 1 : nullptr;
    int *moves_dev = ebuf ? reinterpret_cast<int *>(ebuf + eblocks + 2) : nullptr;
    int *keep_dev = moves_dev ? moves_dev + 1 : nullptr;
    auto energy = [&](const int32_t *lab) -> hipError_t {
        return ginf::launch_energy(GINF_I32, data_dev, table_dev, lab, width, height, num_labels, ebuf, eblocks, en_dev,
                                   st);
    };
    if (e == hipSuccess && result_dev != initial_dev)
        e = hipMemcpyAsync(result_dev, initial_dev, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToDevice, st);
    double best = 0.0;
    if (e == hipSuccess) e = energy(result_dev);
    if (e == hipSuccess) e = hipMemcpyAsync(best_dev, en_dev, sizeof(double), hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ehost, en_dev, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemsetAsync(moves_dev, 0, 2 * sizeof(int), st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) best = *ehost;
    const bool bad_labels = e == hipSuccess && best != best;  // the energy is NaN for a label outside [0, L)
    // Push-relabel needs O(N^2) rounds at worst; this bound only stops a
    // runaway loop.
    const int64_t max_rounds = std::max<int64_t>(4096, 8 * N);
    // The current labelling stays in result_dev: a kept move copies the
    // candidate over it on the device, so the host waits for no energy.
    int moves = 0;
    bool success = true;
    int *moves_host = reinterpret_cast<int *>(ehost + 1);
    for (int t = 0; t < iters && success && e == hipSuccess && !bad_labels; ++t) {
        const int before = moves;
        for (int a = 0; a < num_labels && e == hipSuccess; ++a)
            for (int b = a + 1; b < num_labels && e == hipSuccess; ++b) {
                hipLaunchKernelGGL(k_ab_build, dim3(c.blocks), dim3(256), 0, st, (const int *)data_dev,
                                   (const int *)table_dev, num_labels, result_dev, a, b, g);
                if ((e = hipGetLastError()) != hipSuccess) break;
                if ((e = max_preflow(c, g, max_rounds)) != hipSuccess) break;
                hipLaunchKernelGGL(k_ab_label, dim3(c.blocks), dim3(256), 0, st, result_dev, cand, g, a, b);
                if ((e = energy(cand)) != hipSuccess) break;
                hipLaunchKernelGGL(k_ab_decide, dim3(1), dim3(1), 0, st, en_dev, best_dev, moves_dev, keep_dev);
                hipLaunchKernelGGL(k_ab_commit, dim3(c.blocks), dim3(256), 0, st, keep_dev, cand, result_dev, N);
                e = hipGetLastError();
            }
        // one read of the kept-move count per cycle over the pairs
        if (e == hipSuccess) e = hipMemcpyAsync(moves_host, moves_dev, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess) moves = *moves_host;
        success = moves > before;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (pool) (void)hipFree(pool);
    if (cand) (void)hipFree(cand);
    if (ebuf) (void)hipFree(ebuf);
    if (c.host) (void)hipHostFree(c.host);
    if (ehost) (void)hipHostFree(ehost);
    if (moves_out) *moves_out = moves;
    if (e == hipSuccess && bad_labels) return set_last_error(GINF_ERR_INVALID, "an initial label is outside [0, L)");
    return e == hipSuccess ? GINF_OK : set_last_error(GINF_ERR_HIP, hipGetErrorString(e));
}

int ginf_alpha_beta_host(int width, int height, int num_labels, int dtype, const void *data_host,
                         const void *table_host, const int32_t *initial_host, int iters, int32_t *result_host) {
    if (!data_host || !table_host || !initial_host || !result_host || width < 1 || height < 1 || num_labels < 1)
        return GINF_ERR_INVALID;
    if (dtype != GINF_I32) return ginf::set_last_error(GINF_ERR_UNSUPPORTED, "alpha-beta swap takes int costs only");
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
        rc = ginf_alpha_beta(width, height, num_labels, dtype, dd, dt, (const int32_t *)di, iters, (int32_t *)dr,
                             nullptr, nullptr);
        if (rc == GINF_OK && hipMemcpy(result_host, dr, nl, hipMemcpyDeviceToHost) != hipSuccess) rc = GINF_ERR_HIP;
    }
    for (void *p : {dd, dt, di, dr})
        if (p) (void)hipFree(p);
    return rc;
}

}  // extern "C"
