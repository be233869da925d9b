// Host-side launchers of the BP kernels (bp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "bp_layout.h"

namespace ginf {

// Register-tile shape of a sweep.
//   chunked (default): C labels per lane, G lanes per node split the labels
//   quad:    four lanes per node, one per direction, C = label capacity
//   general: the O(L^2) kernel, one lane per node, C >= L
struct SweepShape {
    int C;
    int G;
    bool general;
    bool quad;
};

// How a node's labels map onto the (plane i, lane q) grid of the state arrays
// (element elem_index(slot, i, q, C, G) of bp_layout.h): label l = q * C + i, or with
// `interleave` l = i * G + q (the quad kernel's 4-label groups).
struct StateLayout {
    int C;
    int G;
    bool interleave;
};

inline StateLayout state_layout(const SweepShape &sh, int L) {
    if (sh.quad) return {(L + 3) / 4, 4, true};
    if (sh.general) return {sh.C, 1, false};
    return {sh.C, sh.G, false};
}

// Label capacities built for the quad kernel.
#define GINF_QUAD_LIST(X) X(16) X(32) X(64) X(96) X(128) X(256)

// Every (C, G) pair built for the O(L) kernel; L <= C * G.
#define GINF_SHAPE_LIST(X) \
    X(16, 1)               \
    X(8, 2)                \
    X(16, 2)               \
    X(8, 4)                \
    X(16, 4)               \
    X(8, 8)                \
    X(24, 4)               \
    X(12, 8)               \
    X(32, 4)               \
    X(16, 8)               \
    X(8, 16)               \
    X(16, 16)              \
    X(16, 32)              \
    X(16, 64)

struct UpsampleArgs {
    const void *coarse[2][kNumDir];
    void *fine[2][kNumDir];
    LevelGeom cg, fg;
    int C, G;
};

// data16: launch the quad instance that reads a.fit_flag and takes the data
// costs in 16-bit storage when it is 0 (quad kernel only).
hipError_t launch_sweep(int dtype, const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                        hipStream_t st, bool data16 = false);
// The first half-sweep of a seeded level with its incoming messages read from
// a prediction (SweepArgs::pred_*, in_base = the prediction; sweep_seeded.hip):
// quad layout, FP32, L % 4 == 0.
hipError_t launch_sweep_seeded(const SweepShape &sh, const SweepArgs &a, bool labels, bool d16, hipStream_t st);
// Per-family launchers (sweep_chunked_*.hip, sweep_quad.hip, sweep_general.hip).
hipError_t launch_sweep_chunked_f32(const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                                    hipStream_t st);
hipError_t launch_sweep_chunked_i32(const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                                    hipStream_t st);
hipError_t launch_sweep_quad(int dtype, const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                             hipStream_t st);
hipError_t launch_sweep_quad16(int dtype, const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                               hipStream_t st);
hipError_t launch_sweep_general(int dtype, const SweepShape &sh, const SweepArgs &a, bool zeroin, bool labels,
                                hipStream_t st);
// Two half-sweeps i (moving colour aA.color) and i + 1 (aB.color) of one
// level in one launch, colour A's messages passed through LDS
// (sweep_pair.inc); quad layout, label capacities 16, 32, 64, 96 and 128.
// flags: 1 = labels of colour A, 2 = labels of colour B, 4 = write colour A's
// outboxes to HBM (no later half-sweep of the level rewrites them).
hipError_t launch_sweep_pair_f32(int lmax, const SweepArgs &aA, const SweepArgs &aB, bool zeroin, int flags,
                                 hipStream_t st);
hipError_t launch_sweep_pair_i32(int lmax, const SweepArgs &aA, const SweepArgs &aB, bool zeroin, int flags,
                                 hipStream_t st);
// Workgroups (tiles) of one fused launch on a level with half-row whp and h rows.
int pair_tiles_f32(int lmax, int whp, int h);
int pair_tiles_i32(int lmax, int whp, int h);
inline bool pair_capacity(int lmax) { return lmax == 16 || lmax == 32 || lmax == 64 || lmax == 96 || lmax == 128; }
hipError_t launch_relayout(int dtype, const void *src, int W, int H, int L, const StateLayout &ly, const LevelGeom &g,
                           void *dst0, void *dst1, hipStream_t st);
// data16: store the level data as 16-bit integers; a value without an exact
// 16-bit form sets *flag (which the caller clears beforehand) and is stored as 0.
// data8 (label-quad layout): the same with 8-bit integers, values in [0, 255].
// only_if: when non-null the kernel does its work only if *only_if != 0 (the
// redo in a wider form after an attempt that did not fit).
hipError_t launch_relayout_pyr(int dtype, const void *src, int W, int H, int L, const StateLayout &ly,
                               const LevelGeom &g0, void *d00, void *d01, bool coarse, const LevelGeom &g1, void *d10,
                               void *d11, hipStream_t st, bool data16 = false, int *flag = nullptr,
                               const int *only_if = nullptr, bool data8 = false);
// The label-quad relayout that also writes the costs min-pooled over 4 x 4
// blocks, times `scale`, as BF16 NHWC [(H+3)/4][(W+3)/4][L] into mp (the
// warm-start network's input); hipErrorNotSupported for other layouts.
hipError_t launch_relayout_quad_mp(int dtype, const void *src, int W, int H, int L, const StateLayout &ly,
                                   const LevelGeom &g0, void *d00, void *d01, bool coarse, const LevelGeom &g1,
                                   void *d10, void *d11, hipStream_t st, bool data16, int *flag, uint16_t *mp,
                                   float scale, bool data8 = false);
// src8 (label-quad layout, with data16): the fine level is stored in 8 bits,
// the coarse one in 16.
hipError_t launch_pyramid(int dtype, const void *f0, const void *f1, const LevelGeom &fg, void *c0, void *c1,
                          const LevelGeom &cg, int C, int G, hipStream_t st, bool data16 = false, int *flag = nullptr,
                          const int *only_if = nullptr, bool src8 = false);
hipError_t launch_upsample(int dtype, const UpsampleArgs &u, hipStream_t st);
// Batched persistent half-sweep (sweep_batch.hip) over instances of label
// capacity lmax (16, 32, 64, 96 or 128), all with L == lmax when `full`:
// args[i] describes instance i for the moving colour; items[k] = (instance,
// first slot, slots).
struct BatchItem {
    int inst;
    int slot0;
    int nslots;
    int pad;
};
// pred: the first half-sweep of a predicted batch (args carry each instance's
// prediction as the incoming-message source; FP32 only).
hipError_t launch_sweep_batch(int dtype, int lmax, bool full, const SweepArgs *args_dev, const BatchItem *items_dev, int nitems,
                              int grid, bool zeroin, bool labels, hipStream_t st, bool pred = false);
// The same with each instance's data costs in 16-bit storage when its
// SweepArgs::fit_flag reads 0 (sweep_batch16.hip).
hipError_t launch_sweep_batch16(int dtype, int lmax, bool full, const SweepArgs *args_dev, const BatchItem *items_dev, int nitems,
                                int grid, bool zeroin, bool labels, hipStream_t st, bool pred = false);

// Seed a level's outboxes from an NHWC prediction (conv_net.hip).
hipError_t launch_inject(const float *pred, void *const msgs[2][4], const LevelGeom &g, int C, int G, bool interleave,
                         int L, int shift, int pw, int ph, float inv_scale, hipStream_t st);

hipError_t launch_block_mean_state(void *const msgs[2][4], const LevelGeom &g, int C, int G, bool interleave, int L,
                                   int f, float scale, float *out, hipStream_t st);

hipError_t launch_fill_u32(void *p, int64_t n, uint32_t v, hipStream_t st);
hipError_t launch_zero_labels(int32_t *labels, int W, int H, int color, hipStream_t st);
hipError_t launch_energy(int dtype, const void *data, const void *table, const int32_t *labels, int W, int H, int L,
                         double *partials, int nblocks, double *out, hipStream_t st);

// Records the message ginf_last_error() returns and passes `code` through.
int set_last_error(int code, const char *msg);

}  // namespace ginf
