/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C restatement of the reference CPU alpha-beta swap exactly as the
 * reference runs it (src/graphcut.cc:9-279, int costs): one sequential
 * push-relabel per label pair whose graph state -- excess, heights and the
 * six residues of every node (S, E, N, W, source, sink) -- lives for the
 * whole decode, and graphCutConstructGraph (graphcut.cc:9-46) rewrites it
 * only for the nodes labelled alpha or beta.  Nodes outside the pair keep
 * what the previous pairs left, and the count of active nodes, the global
 * relabel, the pushes and the reachable set all see it, as in the reference.
 * Every step follows the reference's order: row-major push sweeps with
 * immediate (Gauss-Seidel) excess and residue updates (graphcut.cc:159-177),
 * row-major relabel sweeps reading heights already updated in the sweep
 * (:179-195), ten push/relabel rounds per global relabel (graphcut.h:8,
 * graphcut.cc:237-249), the BFS global relabel with max(height, N+2) for
 * unreachable nodes (:90-140), the sink-reachable set (:47-58, an explicit
 * stack instead of recursion: the set does not depend on the visit order),
 * alpha for reachable pair nodes (:60-72), and the keep-if-cheaper rule on
 * getLabelingCost in int (:265-275).  MIN is (a < b) ? a : b (utils.h:16).
 *
 * The GPU path (csrc/graphcut.hip) starts every pair from a fresh graph
 * instead; this restatement is what its energies are compared with.
 *
 * Pinned (tests/test_oracle_graphcut.py) by an independent Python
 * restatement of graphcut.cc on tiny grids (labels, kept moves), by the
 * two-label case, where every node is in the one pair and the graph state
 * is rebuilt in full, so the result must be the minimum-energy labelling
 * found by brute force and the fresh-graph oracle's labels, and by energy
 * descent.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GR_MIN(a, b) (((a) < (b)) ? (a) : (b))
#define GR_MAX(a, b) (((a) > (b)) ? (a) : (b))
#define GR_SOURCE 4
#define GR_SINK 5
#define GR_ITERS_PER_GLOBAL_RELABEL 10

static const int GX[4] = {0, 1, 0, -1};
static const int GY[4] = {1, 0, -1, 0};
#define GR_OPP(d) ((d) >= 2 ? (d) - 2 : (d) + 2)

typedef struct {
    int W, H, L, N;
    const int *data, *table;
    int *excess, *height, *res, *reach, *stack, *queue, *h2;
} GrState;

/* residue(x, y, k) of Matrix<T>(3, W, H, 6): element x + W*y + W*H*k */
#define RES(s, n, k) ((s)->res[(size_t)(n) + (size_t)(s)->N * (k)])

static int gr_valid(const GrState *s, int x, int y) { return x >= 0 && y >= 0 && x < s->W && y < s->H; }
static int gr_v(const GrState *s, int a, int b) { return s->table[a + s->L * b]; }
static int gr_d(const GrState *s, int n, int l) { return s->data[(size_t)l * s->N + n]; }

/* graphCutConstructGraph (graphcut.cc:9-46) */
static void gr_construct(GrState *s, const int *f, int alpha, int beta) {
    for (int y = 0; y < s->H; y++)
        for (int x = 0; x < s->W; x++) {
            const int n = y * s->W + x;
            if (f[n] != alpha && f[n] != beta) continue;
            int ac = 0, bc = 0;
            for (int d = 0; d < 4; d++) {
                const int nx = x + GX[d], ny = y + GY[d];
                if (gr_valid(s, nx, ny)) {
                    const int fq = f[ny * s->W + nx];
                    if (fq == alpha || fq == beta) {
                        RES(s, n, d) = gr_v(s, alpha, beta);
                    } else {
                        RES(s, n, d) = 0;
                        ac += gr_v(s, alpha, fq);
                        bc += gr_v(s, beta, fq);
                    }
                } else {
                    RES(s, n, d) = 0;
                }
            }
            s->excess[n] = gr_d(s, n, alpha) + ac;
            RES(s, n, GR_SINK) = gr_d(s, n, beta) + bc;
            s->height[n] = 0;
        }
}

/* graphCutCountActive (graphcut.cc:74-88) */
static int gr_count_active(const GrState *s) {
    int active = 0;
    for (int n = 0; n < s->N; n++)
        if (s->excess[n] && s->height[n] < s->N + 2) active++;
    return active;
}

/* graphCutGlobalRelabel (graphcut.cc:90-140) */
static void gr_global_relabel(GrState *s) {
    int qh = 0, qt = 0;
    memset(s->h2, 0, sizeof(int) * (size_t)s->N);
    for (int n = 0; n < s->N; n++)
        if (RES(s, n, GR_SINK)) {
            s->h2[n] = 1;
            s->queue[qt++] = n;
        }
    while (qh < qt) {
        const int n = s->queue[qh++];
        const int x = n % s->W, y = n / s->W;
        for (int d = 0; d < 4; d++) {
            const int nx = x + GX[d], ny = y + GY[d];
            if (!gr_valid(s, nx, ny)) continue;
            const int m = ny * s->W + nx;
            if (RES(s, m, GR_OPP(d)) && s->h2[m] == 0) {
                s->h2[m] = s->h2[n] + 1;
                s->queue[qt++] = m;
            }
        }
    }
    for (int n = 0; n < s->N; n++)
        if (s->h2[n] == 0) s->h2[n] = GR_MAX(s->height[n], s->N + 2);
    memcpy(s->height, s->h2, sizeof(int) * (size_t)s->N);
}

/* graphCutPush + graphCutPushToPixel (graphcut.cc:142-177) */
static void gr_push(GrState *s, int x, int y) {
    const int n = y * s->W + x;
    if (!(s->excess[n] && s->height[n] < s->N + 2)) return;
    if (RES(s, n, GR_SINK) && s->height[n] == 1) {
        const int flow = GR_MIN(RES(s, n, GR_SINK), s->excess[n]);
        s->excess[n] -= flow;
        RES(s, n, GR_SINK) -= flow;
    }
    for (int d = 0; d < 4; d++) {
        /* the reference reads the neighbour's height only when the residue is
         * nonzero, which never happens towards a node outside the grid */
        if (!RES(s, n, d)) continue;
        const int m = (y + GY[d]) * s->W + (x + GX[d]);
        if (s->height[n] == s->height[m] + 1) {
            const int flow = GR_MIN(RES(s, n, d), s->excess[n]);
            s->excess[n] -= flow;
            s->excess[m] += flow;
            RES(s, n, d) -= flow;
            RES(s, m, GR_OPP(d)) += flow;
        }
    }
}

/* graphCutRelabel (graphcut.cc:179-195) */
static void gr_relabel(GrState *s, int x, int y) {
    const int n = y * s->W + x;
    if (!(s->excess[n] && s->height[n] < s->N + 2)) return;
    if (RES(s, n, GR_SINK)) {
        s->height[n] = 1;
    } else {
        int mh = s->N + 2;
        for (int d = 0; d < 4; d++)
            if (RES(s, n, d)) mh = GR_MIN(mh, s->height[(y + GY[d]) * s->W + (x + GX[d])]);
        s->height[n] = mh + 1;
    }
}

/* graphCutFindReachable from every node with sink residue (graphcut.cc:47-58, 253-260) */
static void gr_reachable(GrState *s) {
    memset(s->reach, 0, sizeof(int) * (size_t)s->N);
    for (int start = 0; start < s->N; start++) {
        if (!RES(s, start, GR_SINK) || s->reach[start]) continue;
        int top = 0;
        s->stack[top++] = start;
        s->reach[start] = 1;
        while (top) {
            const int n = s->stack[--top];
            const int x = n % s->W, y = n / s->W;
            for (int d = 0; d < 4; d++) {
                const int nx = x + GX[d], ny = y + GY[d];
                if (!gr_valid(s, nx, ny)) continue;
                const int m = ny * s->W + nx;
                if (RES(s, m, GR_OPP(d)) && !s->reach[m]) {
                    s->reach[m] = 1;
                    s->stack[top++] = m;
                }
            }
        }
    }
}

/* getLabelingCost in int (grid.cc:102-125) */
static int gr_energy(const GrState *s, const int *f) {
    int total = 0;
    for (int y = 1; y < s->H - 1; y++)
        for (int x = 1; x < s->W - 1; x++) {
            const int n = y * s->W + x, l = f[n];
            total += gr_d(s, n, l);
            total += gr_v(s, l, f[n - s->W]);
            total += gr_v(s, l, f[n + 1]);
        }
    return total;
}

/* decodeAlphaBeta (graphcut.cc:197-279).  Returns the number of kept moves,
 * or -1 when a pair's push-relabel loop has not ended after max_loops passes
 * of its while loop (graphcut.cc:228-250; a bound the reference does not
 * have). */
int oracle_alpha_beta_ref(const int *data, const int *table, int W, int H, int L, const int *initial, int iters,
                          int *result, int max_loops) {
    GrState s;
    s.W = W;
    s.H = H;
    s.L = L;
    s.N = W * H;
    s.data = data;
    s.table = table;
    const size_t N = (size_t)s.N;
    s.excess = calloc(N, sizeof(int));
    s.height = calloc(N, sizeof(int));
    s.res = calloc(6 * N, sizeof(int));
    s.reach = calloc(N, sizeof(int));
    s.stack = calloc(N, sizeof(int));
    s.queue = calloc(N, sizeof(int));
    s.h2 = calloc(N, sizeof(int));
    int *nr = calloc(N, sizeof(int));
    memcpy(result, initial, N * sizeof(int));
    memcpy(nr, initial, N * sizeof(int));
    int min_cost = gr_energy(&s, result), kept = 0, failed = 0;
    int success = 1;
    for (int t = 0; t < iters && success && !failed; t++) {
        success = 0;
        for (int alpha = 0; alpha < L && !failed; alpha++)
            for (int beta = alpha + 1; beta < L && !failed; beta++) {
                gr_construct(&s, result, alpha, beta);
                int loops = 0;
                while (1) {
                    if (gr_count_active(&s) == 0) break;
                    if (++loops > max_loops) { failed = 1; break; }
                    gr_global_relabel(&s);
                    for (int i = 0; i < GR_ITERS_PER_GLOBAL_RELABEL; i++) {
                        for (int y = 0; y < H; y++)
                            for (int x = 0; x < W; x++) gr_push(&s, x, y);
                        for (int y = 0; y < H; y++)
                            for (int x = 0; x < W; x++) gr_relabel(&s, x, y);
                    }
                }
                if (failed) break;
                gr_reachable(&s);
                for (size_t n = 0; n < N; n++)
                    if (nr[n] == alpha || nr[n] == beta) nr[n] = s.reach[n] ? alpha : beta;
                const int cost = gr_energy(&s, nr);
                if (cost < min_cost) {
                    min_cost = cost;
                    memcpy(result, nr, N * sizeof(int));
                    success = 1;
                    kept++;
                } else {
                    memcpy(nr, result, N * sizeof(int));
                }
            }
    }
    free(s.excess); free(s.height); free(s.res); free(s.reach); free(s.stack); free(s.queue); free(s.h2); free(nr);
    return failed ? -1 : kept;
}
