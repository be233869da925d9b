/*
 * ginf_bp.h -- C ABI of the MI355X (gfx950) grid min-sum belief propagation
 * library.  Plain pointers and sizes only; no HIP or torch types cross it
 * (streams are passed as void* holding a hipStream_t, NULL = null stream).
 *
 * Every function returns GINF_OK (0) or a positive GINF_ERR_* code and
 * never throws; ginf_last_error() describes the last failure on the
 * calling thread.
 *
 * The BP entry points replace the reference's
 *   template<T> void decodeBp(Grid<T>*, int numIters, Matrix<int>*)          src/include/bp.h:13-14
 *   template<T> void gpuDecodeBp(Grid<T>*, int numIters, Matrix<int>*)       src/include/bp.h:17-18
 *   template<T> void decodeHbp(Grid<T>*, int levels, int iters, Matrix<int>*) src/include/bp.h:21-22
 *   template<T> void gpuDecodeHbp(Grid<T>*, int levels, int iters, Matrix<int>*) src/include/bp.h:25-26
 * implemented in src/bp.cc:86-187 and src/gpu/bp.cu:133-234; the energy
 * replaces Grid<T>::getLabelingCost (src/grid.cc:102-125).  The C++ drop-in
 * templates in bp.h/grid.h/matrix.h are thin wrappers over this ABI.
 */
#ifndef GINF_BP_H_C_ABI
#define GINF_BP_H_C_ABI

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ginf_status {
    GINF_OK = 0,
    GINF_ERR_INVALID = 1,      /* bad argument (sizes, null pointers, enums) */
    GINF_ERR_UNSUPPORTED = 2,  /* valid but not built (e.g. label count too large) */
    GINF_ERR_HIP = 3,          /* HIP runtime failure (message in ginf_last_error) */
    GINF_ERR_NOMEM = 4         /* device allocation failed */
};

/* Element type of costs and messages: the reference instantiates int and float. */
enum ginf_dtype { GINF_F32 = 0, GINF_I32 = 1 };

/* Smoothness models, same values as SmoothnessFunctions (src/include/grid.h:9-14). */
enum ginf_sm_model { GINF_SMOOTH_EXPLICIT = 0, GINF_SMOOTH_POTTS = 1, GINF_SMOOTH_TRUNC_LINEAR = 2, GINF_SMOOTH_TRUNC_QUAD = 3 };

/* Message normalisation: MEAN is the reference (src/bp.cc:44-54); MIN subtracts
 * the message minimum (exact in FP32 for integer-valued costs). */
enum ginf_norm { GINF_NORM_MEAN = 0, GINF_NORM_MIN = 1 };

/* Half-sweep kernel family for the O(L) models (POTTS, TRUNC_LINEAR).
 * CHUNKED: G lanes per node split the labels (exact for integer arithmetic).
 * QUAD: four lanes per node, one per direction, each running the reference's
 * sequential scans and normaliser (exact for every type and normalisation).
 * AUTO: QUAD for FP32 mean normalisation up to 256 labels and for L <= 32,
 * else CHUNKED (up to 1024 labels).  The O(L^2) models (EXPLICIT,
 * TRUNC_QUAD) run one lane per node up to 128 labels and one wavefront per
 * node up to 1024. */
enum ginf_kernel { GINF_KERNEL_AUTO = 0, GINF_KERNEL_CHUNKED = 1, GINF_KERNEL_QUAD = 2 };

typedef struct ginf_bp_desc {
    int width;            /* W >= 1 */
    int height;           /* H >= 1 */
    int num_labels;       /* L >= 1 */
    int dtype;            /* ginf_dtype */
    int sm_model;         /* ginf_sm_model */
    int num_levels;       /* >= 1; 1 = plain loopy BP (decodeBp) */
    int iters_per_level;  /* >= 0 red-black half-sweeps per level */
    int norm;             /* ginf_norm */
    int lanes_per_node;   /* chunked kernel: 0 = default; else 1, 2, 4, ..., 64 lanes share a node's labels */
    int kernel;           /* ginf_kernel */
    int reserved[6];      /* must be zero */
} ginf_bp_desc;

typedef struct ginf_bp_plan_s *ginf_bp_plan;

/* Allocate the device workspace for one problem shape.  sm_table_host is the
 * L x L smoothness table in the reference layout V(a,b) at a + L*b
 * (Matrix<T>(2,L,L), src/grid.cc:8), element type per desc->dtype. */
int ginf_bp_plan_create(const ginf_bp_desc *desc, const void *sm_table_host, ginf_bp_plan *out);

/* Run HBP on device-resident costs data_dev (L x H x W, x fastest, the layout
 * of Grid<T>::getDataCosts()) and write interior labels of labels_dev
 * (H x W int32, x fastest); border labels are left untouched, as the
 * reference leaves them.  Asynchronous on `stream`. */
int ginf_bp_plan_run(ginf_bp_plan plan, const void *data_dev, int32_t *labels_dev, void *stream);

int ginf_bp_plan_destroy(ginf_bp_plan plan);

/* 16- and 8-bit data costs (no reference counterpart).  With the quad
 * kernel a plan stores the pyramid's data costs as 16-bit integers whenever
 * every cost of every level is an integer in [0, 65535], and as T otherwise;
 * levels 0 and 1 are stored in 8 bits when every cost of both is an integer
 * in [0, 255].  The checks and the choice happen on the device during the
 * run (the sweeps read the outcome), every form gives identical results, and
 * a node-label update moves 36 bytes with T costs, 34 with 16-bit and 33
 * with 8-bit costs.  enable: 1 (default) the narrowest form that fits, 2 at
 * most 16-bit storage, 0 always T.  ginf_bp_plan_compact_data reports the
 * last run's form (it waits for the device): 0 T, 1 16-bit, 2 8-bit at
 * levels 0 and 1 and 16-bit above. */
int ginf_bp_plan_set_compact_data(ginf_bp_plan plan, int enable);
int ginf_bp_plan_compact_data(ginf_bp_plan plan, int *last_run_compact);

/* Fused pairs of half-sweeps (no reference counterpart).  With the quad
 * kernel (label capacities 16, 32, 64, 96 and 128), half-sweeps i and i + 1 of
 * a level run in one launch: the first colour's messages reach the second
 * colour through LDS, and go to HBM only where no later half-sweep of the
 * level rewrites them, so an update moves fewer bytes.  Results are
 * identical to the unfused half-sweeps.  mode 1 = on, 0 = off, -1 = the
 * default: on for label capacities 16 and 32, where they measured faster.
 * The plan then holds a second set of its colour-0 outbox arrays (half its
 * message memory again).  ginf_bp_plan_pairs reports whether runs of the
 * plan use them. */
int ginf_bp_plan_set_pairs(ginf_bp_plan plan, int mode);
int ginf_bp_plan_pairs(ginf_bp_plan plan, int *enabled);

/* Bytes of device workspace the plan holds. */
int ginf_bp_plan_workspace_bytes(ginf_bp_plan plan, size_t *bytes);

/* Synchronous host-pointer entry: the drop-in for decodeHbp/gpuDecodeHbp
 * (decodeBp/gpuDecodeBp with num_levels = 1).  Copies costs to the device,
 * runs, copies interior labels back into labels_host.  The reference
 * allocates and frees its state on every call (src/bp.cc:89-179,
 * src/gpu/bp.cu:141-222); here the device state of the last call's problem
 * (descriptor, table, device) is kept and reused by the next call with the
 * same problem, and calls are serialised. */
int ginf_hbp_host(const ginf_bp_desc *desc, const void *data_host, const void *sm_table_host,
                  int32_t *labels_host);
/* Frees the device state ginf_hbp_host keeps (no reference counterpart). */
int ginf_hbp_host_release(void);

/* Labeling energy of Grid::getLabelingCost (src/grid.cc:102-125): interior
 * nodes' D(f) + V(f, f_north) + V(f, f_east), summed in FP64 on the device in
 * a fixed order.  table_dev is the L x L table on the device.  Synchronous:
 * the result is written to *energy_host. */
int ginf_energy_f64(int width, int height, int num_labels, int dtype, const void *data_dev,
                    const void *table_dev, const int32_t *labels_dev, double *energy_host, void *stream);

/* The same energy without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code in tight serving
 * loops, such as per-instance fallback decisions; code that must not
 * synchronise the device through hipMalloc/hipFree).  Its per-block partial
 * sums live in a fixed 8,200-byte device array of the library (1024 partials
 * and the total; each device holds one copy until the process ends), whose
 * size does not depend on the grid, so every grid size is accepted.  All
 * calls on one device share that array: calls that overlap on one device
 * race and must be serialised by the caller.  Synchronous, like
 * ginf_energy_f64. */
int ginf_energy_f64_noalloc(int width, int height, int num_labels, int dtype, const void *data_dev,
                            const void *table_dev, const int32_t *labels_dev, double *energy_host, void *stream);

/* Test hook: copy the outbox messages of pyramid level `level` from the last
 * run into msgs_host in the oracle layout [l][d][y][x] (Matrix<T>(4,w,h,4,L),
 * src/bp.cc:125).  Synchronous. */
int ginf_bp_plan_dump_msgs(ginf_bp_plan plan, int level, void *msgs_host);

/* Profiling hook (no reference counterpart): while enabled, every steady-state
 * half-sweep launch of ginf_bp_plan_run (all but the zero-input first
 * half-sweep of the top level and the two labelling half-sweeps of level 0)
 * is bracketed by a HIP event pair on the run's stream.  ginf_bp_plan_sweep_stats synchronises those events and returns the
 * summed kernel time (ms), the number of timed launches and the node-label
 * updates they performed (updated interior nodes x L); reset != 0 clears. */
int ginf_bp_plan_set_sweep_timing(ginf_bp_plan plan, int enable);
int ginf_bp_plan_sweep_stats(ginf_bp_plan plan, double *total_ms, int64_t *launches, double *node_label_updates,
                             int reset);
/* The timed launches' node-label updates split by level (updates_by_level
 * holds one entry per pyramid level; call before a resetting
 * ginf_bp_plan_sweep_stats), so that a caller can weigh each level's
 * updates by the bytes its data-cost form moves (no reference counterpart). */
int ginf_bp_plan_sweep_level_updates(ginf_bp_plan plan, double *updates_by_level);

/* Node-label updates one ginf_bp_plan_run performs: the sum over levels and
 * half-sweeps of (interior nodes of the moving colour) x L. */
int ginf_bp_plan_updates_per_run(ginf_bp_plan plan, double *node_label_updates);

/* Test hook: copy the data costs of pyramid level `level` from the last run
 * into data_host in the reference layout [l][y][x] (Matrix<T>(3,w,h,L),
 * src/bp.cc:96,109).  Synchronous. */
int ginf_bp_plan_dump_data(ginf_bp_plan plan, int level, void *data_host);

/* Level dimensions of the plan's pyramid (ceil halving, src/bp.cc:101-102). */
int ginf_bp_plan_level_dims(ginf_bp_plan plan, int level, int *w, int *h);

/* Seeded synthetic stereo instance (project-owned generator, no reference
 * counterpart): SplitMix64 stream `seed`; left = 3x3 box-smoothed uint8
 * noise; disparity = background 2 plus `num_rects` rectangles with
 * disparities in [0, dmax]; right(x-d, y) = left(x, y), nearer (larger d)
 * winning, holes filled with noise; cost(l, y, x) = min(|L(x,y) - R(x-l,y)|,
 * trunc), trunc where x - l < 0.  Any output pointer may be NULL.
 * cost_f32 / cost_i32 are L x H x W. */
int ginf_gen_stereo(uint64_t seed, int width, int height, int num_labels, int num_rects, int dmax,
                    int trunc, uint8_t *left, uint8_t *right, int32_t *disp, float *cost_f32,
                    int32_t *cost_i32);

/* Scheduled run (no reference counterpart): like ginf_bp_plan_run, but
 * iters[t] half-sweeps at level t for t <= start_level and levels above
 * start_level skipped.  seeded != 0: level start_level begins from the
 * outboxes already stored there by ginf_bp_plan_seed instead of zeros. */
int ginf_bp_plan_run_schedule(ginf_bp_plan plan, const void *data_dev, int32_t *labels_dev, int start_level,
                              const int *iters, int seeded, void *stream);

/* Seed the outboxes of `level` from a device prediction pred_dev (FP32, NHWC
 * [ph][pw][4L], channel d*L + l, d in S,E,N,W order) at 1/2^shift of the
 * level's resolution: interior node (x, y) takes pred[y>>shift][x>>shift] /
 * scale; border outboxes stay 0.  FP32 plans only. */
int ginf_bp_plan_seed(ginf_bp_plan plan, int level, const float *pred_dev, int pw, int ph, int shift, float scale,
                      void *stream);

/* Level data for a scheduled run, built ahead of it (no reference
 * counterpart): the data-cost pyramid of levels 0..start_level from data_dev,
 * as the runs build it.  With minpool_out non-NULL the same pass over the
 * costs also writes the warm-start network's input: the costs min-pooled over
 * 4 x 4 pixel blocks, times minpool_scale, rounded to BF16 (uint16_t), NHWC
 * [(H+3)/4][(W+3)/4][L] -- the values of ginf_minpool_costs with f = 4. */
int ginf_bp_plan_prepare(ginf_bp_plan plan, const void *data_dev, int start_level, uint16_t *minpool_out,
                         float minpool_scale, void *stream);

/* ginf_bp_plan_run_predicted flag: use the level data of the last
 * ginf_bp_plan_prepare (on the same data) instead of building it again. */
#define GINF_RUN_PREPARED 1

/* Seeded scheduled run without the seeding pass (no reference counterpart):
 * the same results as ginf_bp_plan_seed(plan, start_level, pred_dev, pw, ph,
 * shift, scale) followed by ginf_bp_plan_run_schedule(..., seeded = 1), but
 * where the label-quad kernel runs the level (L % 4 == 0, a 16-byte aligned
 * prediction, iters[start_level] >= 2) the first half-sweep reads its
 * incoming messages from pred_dev itself and no outbox of the seeded level
 * is written before it; other plans fall back to seeding first.  pred_dev
 * must stay valid until the run has executed on `stream`.  flags: 0 or
 * GINF_RUN_PREPARED. */
int ginf_bp_plan_run_predicted(ginf_bp_plan plan, const void *data_dev, int32_t *labels_dev, int start_level,
                               const int *iters, const float *pred_dev, int pw, int ph, int shift, float scale,
                               int flags, void *stream);

/* Learned warm start building blocks (no reference counterpart): BF16 3x3
 * convolutions on MFMA with FP32 accumulation.  Tensors NHWC, weights
 * [Cout][3][3][Cin], bf16 passed as uint16_t.  Cin % 32 == 0, Cout % 32 == 0. */
int ginf_conv3x3_fwd(const uint16_t *x, const uint16_t *w, const float *bias, int N, int H, int W, int Cin, int Cout,
                     int relu, uint16_t *y_bf16, float *y_f32, void *stream);
/* A whole network forward in one call: layer k is ginf_conv3x3_fwd from
 * chans[k] to chans[k+1] channels with weights w[k] and bias b[k] (host
 * arrays of device pointers), ReLU and a bf16 store between layers, FP32
 * output without ReLU after the last.  Layer inputs alternate between act0
 * and act1 (each N*H*W*max hidden channels bf16; unused for one layer).  For
 * inference loops, where one host call per layer costs more than the layer. */
int ginf_convnet_fwd(int layers, const int *chans, const uint16_t *const *w, const float *const *b, const uint16_t *x,
                     int N, int H, int W, uint16_t *act0, uint16_t *act1, float *out, void *stream);
/* dW [Cout][3][3][Cin] and db [Cout] from the layer input x and the
 * pre-activation gradient dz (both bf16 NHWC); deterministic. */
int ginf_conv3x3_bwd_weight(const uint16_t *x, const uint16_t *dz, int N, int H, int W, int Cin, int Cout, float *dw,
                            float *db, void *stream);
/* wf[ci][kh][kw][co] = w[co][2-kh][2-kw][ci]: forward convolution with wf is
 * the input gradient (bwd-data) of w. */
int ginf_conv3x3_flip_weights(const uint16_t *w, int Cout, int Cin, uint16_t *wf, void *stream);
/* dz = bf16(dy * (y > 0)) (y = NULL: no ReLU). */
int ginf_relu_bwd(const float *dy, const uint16_t *y, int64_t n, uint16_t *dz, void *stream);
/* grad = 2 (o - t) / n; per-block squared-error partials (<= 1024). */
int ginf_mse(const float *o, const float *t, int64_t n, float *grad, float *partials, int *nparts, void *stream);
/* Adam step (bias-corrected, step >= 1) on FP32 weights; refreshes the bf16 copy wb when non-NULL. */
int ginf_adam(float *w, const float *g, float *m1, float *m2, uint16_t *wb, int64_t n, float lr, float b1, float b2,
              float eps, int step, float wd, void *stream);
int ginf_to_bf16(const float *s, int64_t n, uint16_t *d, void *stream);
/* Network input: costs [L][H][W] min-pooled over f x f blocks, times scale, bf16 NHWC. */
int ginf_minpool_costs(const void *costs, int dtype, int W, int H, int L, int f, float scale, uint16_t *out,
                       void *stream);
/* Device stereo costs from an image pair (the cost rule of ginf_gen_stereo), FP32 [L][H][W]. */
int ginf_stereo_costs(const uint8_t *left, const uint8_t *right, int W, int H, int L, int trunc, float *out,
                      void *stream);
/* Training target: outboxes [l][d][y][x] averaged over f x f blocks, times scale, FP32 NHWC [.][.][4L]. */
int ginf_block_mean_msgs(const float *msgs, int W, int H, int L, int f, float scale, float *out, void *stream);

/* Training target straight from a plan's state: outboxes of `level` after the
 * last run averaged over f x f node blocks, times scale, FP32 NHWC
 * [ceil(h/f)][ceil(w/f)][4L] with channel d*L + l (FP32 plans). */
int ginf_bp_plan_block_mean_msgs(ginf_bp_plan plan, int level, int f, float scale, float *out_dev, void *stream);

/* Batches (no reference counterpart; the batched workload of the benchmark
 * configurations): n independent single-level MRFs of any sizes and label
 * counts (Potts / truncated linear, L <= 128) sharing dtype, normalisation
 * and iteration count.  Each half-sweep of all instances of one label
 * capacity is one grid-stride launch over a work list; the whole decode is captured once into a HIP
 * graph (re-captured when the data/label pointers change) and replayed.
 * Results equal n separate ginf_bp_plan_run calls with GINF_KERNEL_QUAD. */
typedef struct ginf_batch_s *ginf_batch;
int ginf_batch_create(int n, const ginf_bp_desc *descs, const void *const *tables_host, ginf_batch *out);
int ginf_batch_run(ginf_batch batch, const void *const *data_dev, int32_t *const *labels_dev, void *stream);
/* Tile shape: slots per work item for label capacities 16, 32, 64, 128
 * (multiples of 64) and grid-stride blocks per compute unit. */
int ginf_batch_set_tiles(ginf_batch batch, const int *slots_per_item, int blocks_per_cu);
int ginf_batch_updates_per_run(ginf_batch batch, double *node_label_updates);
/* Seeded batches: with seeded != 0 every instance starts from the outboxes
 * stored by ginf_batch_seed (ginf_bp_plan_seed on the instance) instead of
 * zeros. */
int ginf_batch_set_seeded(ginf_batch batch, int seeded);
/* 16- or 8-bit data costs per instance where they fit, as
 * ginf_bp_plan_set_compact_data (enable 1, the default: the narrowest form
 * that fits; 2: at most 16 bits; 0: T; decided on the device for each
 * instance during the run). */
int ginf_batch_set_compact_data(ginf_batch batch, int enable);
int ginf_batch_seed(ginf_batch batch, int instance, const float *pred_dev, int pw, int ph, int shift, float scale,
                    void *stream);
/* Predictions as the first half-sweep's incoming messages (no reference
 * counterpart): instance i's outboxes are read from pred_dev[i] (FP32 NHWC
 * [ph[i]][pw[i]][4L], 1/2^shift of the instance's resolution, divided by
 * scale[i], border neighbours 0) -- the messages ginf_batch_seed would store,
 * without storing them.  Same results as seeding every instance and running
 * with ginf_batch_set_seeded(batch, 1).  The pointers are read at every
 * ginf_batch_run until the next call; pred_dev == NULL switches it off.
 * FP32 batches with two half-sweeps or more, L % 4 == 0, 16-byte aligned
 * predictions. */
int ginf_batch_set_predictions(ginf_batch batch, const float *const *pred_dev, const int *pw, const int *ph, int shift,
                               const float *scale);
int ginf_batch_dump_msgs(ginf_batch batch, int instance, void *msgs_host);
int ginf_batch_destroy(ginf_batch batch);

/* ---- ICM and the max-prior labelling (decoders adjacent to BP) ----------
 * ICM: decodeIcmSync / gpuDecodeIcm (synchronous, src/icm.cc:5-37,
 * src/gpu/icm.cu:21-83) and decodeIcmAsync (in place, row-major,
 * src/icm.cc:39-62).  Interior nodes move to the lowest-index label of
 * minimal D(l) + V(l,f_S) + V(l,f_E) + V(l,f_N) + V(l,f_W) (T arithmetic,
 * that order); border labels stay as in `initial`.  Device pointers: costs
 * [L][H][W], table V(a,b) at a + L*b, labels [H][W]; result may equal
 * initial.  Stream-ordered. */
enum ginf_icm_mode { GINF_ICM_SYNC = 0, GINF_ICM_ASYNC = 1 };
int ginf_icm(int width, int height, int num_labels, int dtype, const void *data_dev, const void *table_dev,
             const int32_t *initial_dev, int iters, int mode, int32_t *result_dev, void *stream);
/* decodeMaxPrior (src/include/estimators.h:10-25): every node's lowest-index
 * label of minimal data cost. */
int ginf_max_prior(int width, int height, int num_labels, int dtype, const void *data_dev, int32_t *result_dev,
                   void *stream);
/* Alpha-beta swap: decodeAlphaBeta / gpuDecodeAlphaBeta (src/include/graphcut.h:20-26,
 * src/graphcut.cc:194-279), int costs only as in the reference.  For each
 * label pair the swap graph of graphCutConstructGraph (graphcut.cc:9-46) is
 * built from a fresh state, a parallel push-relabel finds a maximum preflow,
 * nodes of the pair that can reach the sink take alpha and the rest beta,
 * and the move is kept when getLabelingCost drops; up to `iters` rounds over
 * all pairs, stopping after a round without a kept move.  *moves_out (may be
 * NULL) receives the number of kept moves.  Synchronous on `stream`. */
int ginf_alpha_beta(int width, int height, int num_labels, int dtype, const void *data_dev, const void *table_dev,
                    const int32_t *initial_dev, int iters, int32_t *result_dev, int *moves_out, void *stream);
int ginf_alpha_beta_host(int width, int height, int num_labels, int dtype, const void *data_host,
                         const void *table_host, const int32_t *initial_host, int iters, int32_t *result_host);
/* Synchronous host-pointer forms (used by include/icm.h and estimators.h). */
int ginf_icm_host(int width, int height, int num_labels, int dtype, const void *data_host, const void *table_host,
                  const int32_t *initial_host, int iters, int mode, int32_t *result_host);
int ginf_max_prior_host(int width, int height, int num_labels, int dtype, const void *data_host,
                        int32_t *result_host);

/* HBM calibration probe (no reference counterpart): reads `reads` and
 * writes `writes` float4 streams of n_vec4 elements each ((reads, writes) =
 * (1, 1) or (5, 4), the half-sweep's stream shape); in_dev holds reads *
 * n_vec4 float4, out_dev writes * n_vec4.  Stream-ordered. */
int ginf_stream_probe(int reads, int writes, const void *in_dev, void *out_dev, int64_t n_vec4, int blocks,
                      void *stream);

const char *ginf_last_error(void);

/* Library version string. */
const char *ginf_version(void);

#ifdef __cplusplus
}
#endif

#endif
