/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C restatement of the CPU min-sum belief propagation of
 * GridInfCuda (src/bp.cc, with the energy of src/grid.cc:102-125 and the
 * Matrix layout of src/matrix.cc).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the reference results are
 * checked against or as the timed CPU baseline; the HIP library never calls it.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   oracle_pyramid      brute-force child sums, ceil(w/2) dims (numpy)
 *   oracle_upsample     closed-form parent copy (numpy)
 *   send_msg O(L) path  brute-force min_j(h_j + V(i,j)) on random integer
 *                       vectors; Potts identity; mean/min normalisation
 *                       invariants (both through oracle_hbp on 3x3 grids)
 *   send_msg O(L^2)     brute force on 3x3 grids (EXPLICIT / TRUNC_QUAD)
 *   oracle_half_sweep   colour parity known answers (iters=1: nodes with
 *                       x+y even keep label 0), an independent dict-based
 *                       Python message-passing restatement on tiny grids,
 *                       and exact Viterbi MAP on chain grids (H=3)
 *   oracle_labels       lowest-index argmin; 3x3 grid = decodeMaxPrior
 *   oracle_hbp          HBP on H=3 equals plain BP (coarse levels have no
 *                       interior); level handling vs the numpy restatement
 *   oracle_energy*      brute-force FP64/numpy sums
 * The reference itself cannot be compiled here (src/include/bp.h:5 pulls in
 * src/include/gpu.h, which includes <cuda.h>/<cuda_runtime.h>; the image has
 * no CUDA toolkit), so no oracle/_ref build exists.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORACLE_SM_EXPLICIT = 0, ORACLE_SM_POTTS = 1, ORACLE_SM_TRUNC_LINEAR = 2, ORACLE_SM_TRUNC_QUAD = 3 };
enum { ORACLE_NORM_MEAN = 0, ORACLE_NORM_MIN = 1 };

/* Direction order S, E, N, W and their offsets (src/include/utils.h:26-37). */
static const int DIR_X[4] = {0, 1, 0, -1};
static const int DIR_Y[4] = {1, 0, -1, 0};
#define OPP_DIR(d) ((d) >= 2 ? (d) - 2 : (d) + 2)
/* The reference MIN keeps the second operand on ties (src/include/utils.h:16). */
#define OMIN(a, b) (((a) < (b)) ? (a) : (b))

#define OT float
#define OS f32
#include "bp_oracle_t.h"
#undef OT
#undef OS

#define OT int
#define OS i32
#include "bp_oracle_t.h"
#undef OT
#undef OS

/* ---- alpha-beta swap (int), after src/graphcut.cc:9-279 ------------------
 * The graph of one label pair is built as graphCutConstructGraph does
 * (graphcut.cc:9-46): for every node labelled alpha or beta, n-links of
 * capacity V(alpha,beta) to neighbours in the pair, a source t-link
 * D(alpha) + sum V(alpha, f_q) and a sink t-link D(beta) + sum V(beta, f_q)
 * over the neighbours q outside the pair.  Unlike the reference, every pair
 * starts from a fresh graph (the reference keeps the excess, residues and
 * heights of nodes outside the pair from the previous pair).  The maximum
 * flow comes from BFS augmenting paths (Edmonds-Karp), not push-relabel;
 * the nodes that can still reach the sink in the residual graph -- the same
 * set for every maximum flow -- take alpha, the rest of the pair beta
 * (graphCutLabelFromFlow, graphcut.cc:66-79), and the move is kept when it
 * lowers getLabelingCost (graphcut.cc:263-272). */
static void ab_build(const int *data, const int *table, int W, int H, int L, const int *f, int alpha, int beta,
                     int *src, int *rs, int *r, int *in) {
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t n = (size_t)y * W + x;
            int fl = f[n];
            in[n] = fl == alpha || fl == beta;
            src[n] = rs[n] = 0;
            for (int d = 0; d < 4; d++) r[4 * n + d] = 0;
            if (!in[n]) continue;
            int ac = 0, bc = 0;
            for (int d = 0; d < 4; d++) {
                int nx = x + DIR_X[d], ny = y + DIR_Y[d];
                if (nx < 0 || ny < 0 || nx >= W || ny >= H) continue;
                int fq = f[(size_t)ny * W + nx];
                if (fq == alpha || fq == beta) r[4 * n + d] = table[alpha + L * beta];
                else { ac += table[alpha + L * fq]; bc += table[beta + L * fq]; }
            }
            src[n] = data[(size_t)alpha * W * H + n] + ac;
            rs[n] = data[(size_t)beta * W * H + n] + bc;
        }
}

/* Edmonds-Karp on the grid graph; afterwards reach[n] = node n can reach the
 * sink in the residual graph. */
static void ab_maxflow(int W, int H, int *src, int *rs, int *r, int *reach, int *prev, int *queue) {
    size_t N = (size_t)W * H;
    for (;;) {
        /* BFS from the source: prev[n] = -1 unvisited, -2 entered from the source, else 4*u + d */
        for (size_t n = 0; n < N; n++) prev[n] = -1;
        size_t qh = 0, qt = 0;
        for (size_t n = 0; n < N; n++)
            if (src[n] > 0) { prev[n] = -2; queue[qt++] = (int)n; }
        int last = -1;
        while (qh < qt && last < 0) {
            int u = queue[qh++];
            if (rs[u] > 0) { last = u; break; }
            int ux = u % W, uy = u / W;
            for (int d = 0; d < 4; d++) {
                int vx = ux + DIR_X[d], vy = uy + DIR_Y[d];
                if (vx < 0 || vy < 0 || vx >= W || vy >= H) continue;
                int v = vy * W + vx;
                if (r[4 * u + d] > 0 && prev[v] == -1) { prev[v] = 4 * u + d; queue[qt++] = v; }
            }
        }
        if (last < 0) break;
        int delta = rs[last];
        for (int v = last; prev[v] != -2; v = prev[v] / 4) delta = OMIN(delta, r[prev[v]]);
        { int v = last; while (prev[v] != -2) v = prev[v] / 4; delta = OMIN(delta, src[v]); }
        rs[last] -= delta;
        int v = last;
        while (prev[v] != -2) {
            int e = prev[v], u = e / 4, d = e % 4;
            r[e] -= delta;
            r[4 * v + OPP_DIR(d)] += delta;
            v = u;
        }
        src[v] -= delta;
    }
    /* backward BFS from the sink over residual edges v -> u */
    size_t qh = 0, qt = 0;
    for (size_t n = 0; n < N; n++) {
        reach[n] = rs[n] > 0;
        if (reach[n]) queue[qt++] = (int)n;
    }
    while (qh < qt) {
        int u = queue[qh++];
        int ux = u % W, uy = u / W;
        for (int d = 0; d < 4; d++) {
            int vx = ux + DIR_X[d], vy = uy + DIR_Y[d];
            if (vx < 0 || vy < 0 || vx >= W || vy >= H) continue;
            int v = vy * W + vx;
            if (!reach[v] && r[4 * v + OPP_DIR(d)] > 0) { reach[v] = 1; queue[qt++] = v; }
        }
    }
}

/* One alpha-beta swap move of labelling f (in place). */
void oracle_ab_move(const int *data, const int *table, int W, int H, int L, int *f, int alpha, int beta) {
    size_t N = (size_t)W * H;
    int *src = malloc(N * sizeof(int)), *rs = malloc(N * sizeof(int)), *r = malloc(4 * N * sizeof(int));
    int *in = malloc(N * sizeof(int)), *reach = malloc(N * sizeof(int)), *prev = malloc(N * sizeof(int));
    int *queue = malloc(N * sizeof(int));
    ab_build(data, table, W, H, L, f, alpha, beta, src, rs, r, in);
    ab_maxflow(W, H, src, rs, r, reach, prev, queue);
    for (size_t n = 0; n < N; n++)
        if (in[n]) f[n] = reach[n] ? alpha : beta;
    free(src); free(rs); free(r); free(in); free(reach); free(prev); free(queue);
}

/* decodeAlphaBeta (graphcut.cc:194-279) with fresh per-pair graphs.
 * Returns the number of accepted moves. */
int oracle_alpha_beta(const int *data, const int *table, int W, int H, int L, const int *initial, int iters,
                      int *result) {
    size_t N = (size_t)W * H;
    int *cand = malloc(N * sizeof(int));
    memcpy(result, initial, N * sizeof(int));
    int best = oracle_energy_i32(data, table, result, W, H, L), accepted = 0;
    int success = 1;
    for (int t = 0; t < iters && success; t++) {
        success = 0;
        for (int a = 0; a < L; a++)
            for (int b = a + 1; b < L; b++) {
                memcpy(cand, result, N * sizeof(int));
                oracle_ab_move(data, table, W, H, L, cand, a, b);
                int e = oracle_energy_i32(data, table, cand, W, H, L);
                if (e < best) {
                    best = e;
                    memcpy(result, cand, N * sizeof(int));
                    success = 1;
                    accepted++;
                }
            }
    }
    free(cand);
    return accepted;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
