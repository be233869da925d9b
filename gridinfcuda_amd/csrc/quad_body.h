// Per-node body of the quad half-sweep (one lane per node and direction),
// shared by the single-grid kernel (sweep_quad.hip) and the batched
// persistent kernel (sweep_batch.hip).
#pragma once

#include "sweep_common.h"

// Cache policy of the streaming buffer accesses (0 = default; 2 = non-temporal).
// Outboxes are not read again within the half-sweep, so they are stored
// non-temporally (measured faster on 1920x1080x128); loads keep the default
// (1920x1080x128 HBP per decode: default 11.41-11.44 ms, non-temporal loads
// 11.73 ms, policy 1 loads 11.41 ms).
#ifndef GINF_QUAD_LOAD_POLICY
#define GINF_QUAD_LOAD_POLICY 0
#endif
#ifndef GINF_QUAD_STORE_POLICY
#define GINF_QUAD_STORE_POLICY 2
#endif

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
//   LMAX  register capacity in labels; FULL: L == LMAX (no label guards)
//   D16   instance for plans with 16-bit data storage: *a.fit_flag == 0 means
//         the costs were stored as 16-bit unsigned integers (exact: every cost
//         of every level is an integer in [0, 65535]) and a lane loads its 4
//         labels' costs as 8 bytes; otherwise they were stored in T.  The flag
//         is uniform, so the branch never diverges.
//   D8    (with D16) also the 8-bit form: where a.fit8_flag is set and reads
//         0 as well, this level's costs are integers in [0, 255] stored in
//         bytes, and a lane loads its 4 labels' costs as 4 bytes.
// Layout: label quads in slot tiles, element (l, slot) at
// elem_index(slot, l/4, l%4, C, 4): one lane moves 4 labels (16 B) per
// access, 16 consecutive nodes of a wave cover 256 contiguous bytes of every
// array, and the next label quad of the same nodes lies kTileSlots * 16 B
// further on.
// Offsets are 32-bit and
// carried entirely in the VGPR offset (soffset = 0): with the plane offset in
// soffset, the compiler reused SGPRs that a v_cmp in the labelling code had
// just written, and the memory instruction then intermittently read the wrong
// offset (nondeterministic message corruption, measured on gfx950).
//   PRED  the incoming messages come from a network prediction at
//         1/2^pred_shift of the level's resolution instead of the partner
//         colour's outboxes (first half-sweep of a seeded level): each value is
//         the prediction times pred_inv_scale, 0 from a border neighbour -- the
//         outboxes ginf_bp_plan_seed would have stored, read where they come
//         from (FP32 only; a lane's label quad is 16 contiguous bytes of the
//         prediction row, L % 4 == 0).
// LABELS with LMAX <= 128: the quad writes the node's beliefs to its row of
// the block's LDS (`brow`, LMAX floats) inside the label loop, and
// the quad takes the minimum from there after the outboxes are stored, when
// the message registers are free; lanes hold no running minimum in the loop
// (that kept the labelling instance at two waves per SIMD instead of three).
// LMAX = 256 keeps the running per-lane minimum.
template <bool LABELS, int LMAX>
constexpr bool kLdsLabels = LABELS && LMAX <= 128 && LMAX % 16 == 0;

template <typename T, int LMAX, bool ZEROIN, bool LABELS, bool FULL, bool use16, bool PRED = false, bool use8 = false>
__device__ __forceinline__ void quad_node_impl(const SweepArgs &a, const int slot, const int d, T *brow) {
    typedef T v4 __attribute__((ext_vector_type(4)));
    constexpr int LQ = LMAX / 4;
    if (slot >= a.nslots) return;
    const int y = slot / a.whp;
    const int x = 2 * (slot - y * a.whp) + ((y + a.color) & 1);
    if (y < 1 || y > a.h - 2 || x < 1 || x > a.w - 2) return;  // whole quads leave together
    const int L = FULL ? LMAX : a.num_labels;
    const T s = Num<T>::s(a);
    const T tau = Num<T>::tau(a);

    const __amdgpu_buffer_rsrc_t drs = __builtin_amdgcn_make_buffer_rsrc((void *)a.data, 0, a.data_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t irs = __builtin_amdgcn_make_buffer_rsrc((void *)a.in_base, 0, a.in_span, 0x00020000);
    const __amdgpu_buffer_rsrc_t ors = __builtin_amdgcn_make_buffer_rsrc(a.out_base, 0, 4u * a.arr_bytes, 0x00020000);
    const uint32_t self_off = (uint32_t)(slot >> kTileShift) * a.tile_bytes + (uint32_t)(slot & (kTileSlots - 1)) * 16u;
    // 16-bit storage: every offset halves; 8-bit: a quarter
    const uint32_t data_off = use8 ? self_off / 4u : use16 ? self_off / 2u : self_off;
    const uint32_t data_plane = use8 ? a.data_plane_bytes / 2u : use16 ? a.data_plane_bytes : a.plane_bytes;
    uint32_t in_off = 0;
    bool in_live = true;
    if constexpr (PRED) {
        const int nx = x + kDX[d], ny = y + kDY[d];
        const int od = (d + 2) & 3;
        in_live = nx >= 1 && ny >= 1 && nx <= a.w - 2 && ny <= a.h - 2;
        const int X = min(nx >> a.pred_shift, a.pred_w - 1), Y = min(ny >> a.pred_shift, a.pred_h - 1);
        in_off = ((uint32_t)((Y * a.pred_w + X) * 4 + od) * (uint32_t)L) * 4u;
    } else if (!ZEROIN) {
        int nx = x + kDX[d], ny = y + kDY[d];
        int csel;
        if (a.src == kSrcCoarse) { nx >>= 1; ny >>= 1; csel = (nx + ny) & 1; }
        else csel = 0;  // in_base already points at the partner colour's arrays
        const int od = (d + 2) & 3;
        const uint32_t ns = (uint32_t)(ny * a.in_whp + (nx >> 1));
        in_off = (uint32_t)(csel * 4 + od) * a.in_arr_bytes + (ns >> kTileShift) * a.tile_bytes +
                 (ns & (kTileSlots - 1)) * 16u;
    }
    const uint32_t out_off = (uint32_t)d * a.arr_bytes + self_off;

    T f[LMAX];
    T mu = T(0);
    // Labels: lane d tracks the belief minimum over labels l = 4j + d only
    // (strict <, so the lowest such label wins); the quad combines the four
    // residues at the end.
    T best = T(0);
    int arg = -1;
#pragma unroll
    for (int j = 0; j < LQ; ++j) {
        if (FULL || 4 * j < L) {
            v4 m4 = v4(T(0));
            if (!ZEROIN)
                m4 = __builtin_bit_cast(v4, __builtin_amdgcn_raw_buffer_load_b128(irs, in_off + (uint32_t)j * a.in_plane_bytes, 0, GINF_QUAD_LOAD_POLICY));
            if constexpr (PRED) {
#pragma unroll
                for (int r = 0; r < 4; ++r) m4[r] = in_live ? m4[r] * a.pred_inv_scale : T(0);
            }
            v4 d4;
            if constexpr (use8) {
                const uint32_t h = __builtin_amdgcn_raw_buffer_load_b32(drs, data_off + (uint32_t)j * data_plane, 0, GINF_QUAD_LOAD_POLICY);
                d4[0] = T(h & 0xFFu);
                d4[1] = T((h >> 8) & 0xFFu);
                d4[2] = T((h >> 16) & 0xFFu);
                d4[3] = T(h >> 24);
            } else if constexpr (use16) {
                const auto h = __builtin_amdgcn_raw_buffer_load_b64(drs, data_off + (uint32_t)j * data_plane, 0, GINF_QUAD_LOAD_POLICY);
                d4[0] = T(h[0] & 0xFFFFu);
                d4[1] = T(h[0] >> 16);
                d4[2] = T(h[1] & 0xFFFFu);
                d4[3] = T(h[1] >> 16);
            } else {
                d4 = __builtin_bit_cast(v4, __builtin_amdgcn_raw_buffer_load_b128(drs, data_off + (uint32_t)j * data_plane, 0, GINF_QUAD_LOAD_POLICY));
            }
            // Beliefs of the four labels, D + in_S + in_E + in_N + in_W added
            // left to right (a + b == b + a exactly, so the broadcast operand
            // may come first).  The four chains are interleaved so that no DPP
            // read waits on the add just before it.
            v4 b4;
#pragma unroll
            for (int r = 0; r < 4; ++r) b4[r] = quad_bcast<0>(m4[r]) + d4[r];
#pragma unroll
            for (int r = 0; r < 4; ++r) b4[r] = quad_bcast<1>(m4[r]) + b4[r];
#pragma unroll
            for (int r = 0; r < 4; ++r) b4[r] = quad_bcast<2>(m4[r]) + b4[r];
#pragma unroll
            for (int r = 0; r < 4; ++r) b4[r] = quad_bcast<3>(m4[r]) + b4[r];
            // every lane of the quad holds the same b4 and stores it (a
            // store by lane 0 alone would split the unrolled loop into
            // branches and stop the loads from being issued ahead)
            if constexpr (kLdsLabels<LABELS, LMAX>) *reinterpret_cast<v4 *>(brow + 4 * j) = b4;
            T bsel = T(0);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int l = 4 * j + r;
                if (FULL || l < L) {
                    const T m = m4[r];
                    const T b = b4[r];
                    if (LABELS && !kLdsLabels<LABELS, LMAX> && r == d) bsel = b;
                    const T h = b - m;
                    if (l == 0) { f[0] = h; mu = h; }
                    else { f[l] = tmin(h, f[l - 1] + s); mu = tmin(mu, h); }
                }
            }
            if (LABELS && !kLdsLabels<LABELS, LMAX> && (FULL || 4 * j + d < L)) {
                if (arg < 0 || bsel < best) { best = bsel; arg = 4 * j + d; }
            }
        }
    }
#pragma unroll
    for (int l = LMAX - 2; l >= 0; --l)
        if (FULL || l <= L - 2) f[l] = tmin(f[l], f[l + 1] + s);
    const T cap = mu + tau;
    T acc = T(0), mn = T(0);
#pragma unroll
    for (int l = 0; l < LMAX; ++l) {
        if (FULL || l < L) {
            f[l] = tmin(f[l], cap);
            acc = acc + f[l];
            mn = (l == 0) ? f[0] : tmin(mn, f[l]);
        }
    }
    const T nu = (a.norm == 0) ? acc / T(L) : mn;
#pragma unroll
    for (int j = 0; j < LQ; ++j) {
        if (FULL || 4 * j < L) {
            v4 o;
#pragma unroll
            for (int r = 0; r < 4; ++r) o[r] = (FULL || 4 * j + r < L) ? f[4 * j + r] - nu : T(0);
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, o),
                                                   ors, out_off + (uint32_t)j * a.plane_bytes, 0, GINF_QUAD_STORE_POLICY);
        }
    }
    if constexpr (kLdsLabels<LABELS, LMAX>) {
        // lane d: the lowest-index minimum over labels [d LMAX/4, (d+1) LMAX/4)
        constexpr int R = LMAX / 4;
#pragma unroll
        for (int k = 0; k < R / 4; ++k) {
            const v4 b4 = *reinterpret_cast<const v4 *>(brow + d * R + 4 * k);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int l = d * R + 4 * k + r;
                if ((FULL || l < L) && (arg < 0 || b4[r] < best)) { best = b4[r]; arg = l; }
            }
        }
    }
    if (LABELS) {
        // Combine the residues: smaller belief wins, ties go to the lower label.
#pragma unroll
        for (int k = 1; k < 4; k <<= 1) {
            const T ob = __shfl_xor(best, k, 4);
            const int oa = __shfl_xor(arg, k, 4);
            const bool take = oa >= 0 && (arg < 0 || ob < best || (!(best < ob) && oa < arg));
            if (take) { best = ob; arg = oa; }
        }
        if (d == 0) a.labels[(int64_t)y * a.label_w + x] = arg;
    }
}

// The two storage forms are separate straight-line bodies; the uniform flag
// picks one per wave (a branch inside the unrolled label loop would stop the
// compiler from issuing the loads ahead).
// The belief rows of the labelling instances are declared here, once per
// kernel, and shared by the storage-form bodies; node threadIdx.x / 4 of a
// block (at most 256 threads, 64 nodes) owns row threadIdx.x / 4.
template <typename T, int LMAX, bool ZEROIN, bool LABELS, bool FULL, bool D16 = false, bool PRED = false,
          bool D8 = false>
__device__ __forceinline__ void quad_node(const SweepArgs &a, const int slot, const int d) {
    T *brow = nullptr;
    if constexpr (kLdsLabels<LABELS, LMAX>) {
        // row pitch LMAX + 4 floats: the 16 lane-0 writes of a wave start on distinct bank quads
        __shared__ __attribute__((aligned(16))) T beliefs[64 * (LMAX + 4)];
        brow = beliefs + (threadIdx.x >> 2) * (LMAX + 4);
    }
    if constexpr (D16) {
        if (__builtin_amdgcn_readfirstlane(*a.fit_flag) == 0) {
            if constexpr (D8) {
                if (a.fit8_flag && __builtin_amdgcn_readfirstlane(*a.fit8_flag) == 0) {
                    quad_node_impl<T, LMAX, ZEROIN, LABELS, FULL, true, PRED, true>(a, slot, d, brow);
                    return;
                }
            }
            quad_node_impl<T, LMAX, ZEROIN, LABELS, FULL, true, PRED>(a, slot, d, brow);
            return;
        }
    }
    quad_node_impl<T, LMAX, ZEROIN, LABELS, FULL, false, PRED>(a, slot, d, brow);
}

}  // namespace
}  // namespace ginf
