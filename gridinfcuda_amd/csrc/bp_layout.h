// Device-side data layout of the grid BP state and the kernel argument blocks.
//
// A level of w x h nodes is split by colour c = (x + y) & 1 (the red-black
// classes of the reference schedule, src/bp.cc:149-150).  Node (x, y) of
// colour c lives in slot n = y * whp + (x >> 1) of that colour's arrays,
// where whp >= ceil(w / 2) is the padded half-row.  A node's L labels are
// spread over G lanes ("lanes per node"): lane q owns labels q*C .. q*C+C-1
// (register plane i holds label q*C + i, or 4i + q for the label-quad layout).
// Slots are grouped in tiles of kTileSlots (S = 256) slots; an array of C
// planes stores tile-major, then plane, then slot within the tile, then lane:
//   element(n, i, q) = ((n / S) * C + i) * S * G + (n % S) * G + q.
// One wavefront reading plane i of 64/G consecutive nodes touches 64
// consecutive elements (256 B) of every array, and a wave's successive planes
// follow each other at S*G elements, so a half-sweep reads each array as a
// few long runs instead of C runs one plane (nplane slots) apart.  Measured on
// the 1920x1080x128 decode (5 levels x 10 half-sweeps): 11.61 ms with S = 256,
// 11.75 ms with S = 64, 11.84 ms with S = 1024, 12.17 ms plane-major.
//
// Per level and colour there is one data-cost array and four outbox arrays
// (directions S, E, N, W as in src/include/utils.h:26-37).  Outbox (x,y,d) is
// the message node (x,y) sends towards its neighbour in direction d.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ginf {

constexpr int kNumDir = 4;
constexpr int kSrcSame = 0;    // incoming messages from this level, other colour
constexpr int kSrcCoarse = 1;  // incoming messages read through the parent at level+1

// log2 of the slots per tile.
#ifndef GINF_TILE_SHIFT
#define GINF_TILE_SHIFT 8
#endif
constexpr int kTileShift = GINF_TILE_SHIFT;
constexpr int kTileSlots = 1 << kTileShift;

struct LevelGeom {
    int w, h;        // nodes
    int whp;         // padded half-row (slots per row per colour)
    int nslots;      // h * whp
    int64_t nplane;  // slots per label-register plane (nslots rounded up to whole tiles)
};

// Element of (slot n, plane 0, lane 0) in an array of C planes and G lanes;
// plane i of that slot is kTileSlots * G elements further on.
__host__ __device__ inline int64_t slot_elem(int64_t n, int C, int G) {
    return ((n >> kTileShift) * C * kTileSlots + (n & (kTileSlots - 1))) * G;
}
__host__ __device__ inline int64_t elem_index(int64_t n, int i, int q, int C, int G) {
    return slot_elem(n, C, G) + (int64_t)i * kTileSlots * G + q;
}

// Arguments of one red-black half-sweep launch (linear and general models).
struct SweepArgs {
    const void *data;        // data costs of the moving colour
    void *out[kNumDir];      // outboxes of the moving colour
    const void *in[2][kNumDir];  // [colour][dir] outbox arrays incoming messages are read from
    int64_t stride;          // elements between register planes of a slot (kTileSlots * G)
    int64_t in_stride;       // same for the arrays in `in`
    int64_t tile;            // elements between slot tiles (C * kTileSlots * G), every level of a plan
    int w, h, whp, nslots;   // this level
    int in_whp;              // half-row of the `in` level
    int color;               // moving colour: nodes with (x + y) & 1 == color
    int src;                 // kSrcSame / kSrcCoarse
    int num_labels;          // L
    int norm;                // 0 mean, 1 min
    int32_t *labels;         // W x H result (level 0), written when the kernel is built for it
    int label_w;
    float s_f, tau_f;        // s = V(0,1), tau = V(0,L-1) (src/bp.cc:35,43)
    int s_i, tau_i;
    const void *table;       // L x L table (general model)
    // Byte-addressed view used by the quad kernel (buffer loads with 32-bit
    // offsets): the four outbox arrays of one colour are contiguous,
    // arr_bytes apart, and colour 1's follow colour 0's.
    const void *in_base;     // outbox array [colour 0 or the moving colour's partner][S] of the source level
    void *out_base;          // outbox array [color][S] of this level
    uint32_t arr_bytes;      // bytes between consecutive outbox arrays, this level
    uint32_t in_arr_bytes;   // same, source level
    uint32_t plane_bytes;    // bytes between register planes of a slot (kTileSlots * G * 4)
    uint32_t in_plane_bytes; // same, source level
    uint32_t tile_bytes;     // bytes between slot tiles (C * kTileSlots * G * 4), every level
    uint32_t data_bytes;     // bytes of the data-cost array of the moving colour
    uint32_t in_span;        // bytes addressable from in_base
    int lmax;                // label capacity of the quad kernel instance (batched launches)
    uint32_t data_plane_bytes;  // bytes between data-cost planes in 16-bit storage (half of plane_bytes;
                                // 16-bit element offsets are the T offsets halved)
    const int *fit_flag;     // 16-bit data path: 0 = the level data is stored in 16 bits, else in T
    const int *fit8_flag;    // levels 0 and 1 of a plan whose run tried 8-bit storage: with *fit_flag
                             // 0, *fit8_flag 0 = this level's data is stored in 8 bits (element
                             // offsets a quarter of T's), else in 16; null = never 8 bits
    // Prediction-sourced first half-sweep of a seeded level (quad kernel, FP32):
    // the incoming message from neighbour n in direction od is
    // pred[min(ny >> pred_shift, pred_h - 1)][min(nx >> pred_shift, pred_w - 1)][od * L + l]
    // * pred_inv_scale for an interior n, 0 for a border n.  in_base then
    // points at the prediction and in_span is its size in bytes.
    int pred_w, pred_h, pred_shift;
    float pred_inv_scale;
};

}  // namespace ginf
