// C ABI of the grid BP library: plans (device workspace per problem shape),
// the HBP driver, the synchronous host-pointer entry, the FP64 energy and
// test hooks.  The driver follows decodeHbp (reference src/bp.cc:86-180):
// data-cost pyramid, then levels coarse to fine, each starting from the
// outbox messages of the level above and running `iters` red-black
// half-sweeps, then the stale-belief labelling of the finest level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "bp_kernels.h"
#include "ginf_bp.h"

namespace {

thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

}  // namespace

// Shared with the other translation units of the library (declared in bp_kernels.h).
int ginf::set_last_error(int code, const char *msg) {
    g_last_error = msg;
    return code;
}

namespace {

#define GINF_HIP_TRY(expr)                                                                                   \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return set_error(GINF_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                                      \
    } while (0)

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Layout knob.  GINF_ARRAY_STAGGER: extra bytes between a level's state
// arrays (0 = packed; with plane-major arrays, 4, 8 and 68 KiB staggers were
// all slower on the 1920x1080x128 half-sweep).
#ifndef GINF_ARRAY_STAGGER
#define GINF_ARRAY_STAGGER 0
#endif

// Default register tile per label count for the O(L) kernel: the fewest
// lanes per node that keep C <= 16 labels per lane.
bool pick_linear_shape(int L, int lanes, ginf::SweepShape &sh) {
    struct P { int C, G; };
    static const P pairs[] = {
#define GINF_PAIR(CC, GG) {CC, GG},
        GINF_SHAPE_LIST(GINF_PAIR)
#undef GINF_PAIR
    };
    if (lanes == 0) {
        int g = L <= 16 ? 1 : L <= 32 ? 2 : L <= 64 ? 4 : L <= 128 ? 8 : L <= 256 ? 16 : L <= 512 ? 32 : L <= 1024 ? 64 : 0;
        if (g == 0) return false;
        if (L > 64 && L <= 96) { sh = {12, 8, false, false}; return true; }
        sh = {16, g, false, false};
        return true;
    }
    int best = -1;
    for (int i = 0; i < (int)(sizeof pairs / sizeof pairs[0]); ++i)
        if (pairs[i].G == lanes && pairs[i].C * pairs[i].G >= L && (best < 0 || pairs[i].C < pairs[best].C)) best = i;
    if (best < 0) return false;
    sh = {pairs[best].C, pairs[best].G, false, false};
    return true;
}

bool pick_quad_shape(int L, ginf::SweepShape &sh) {
    static const int caps[] = {
#define GINF_CAP(LM) LM,
        GINF_QUAD_LIST(GINF_CAP)
#undef GINF_CAP
    };
    for (int c : caps)
        if (L <= c) { sh = {c, 4, false, true}; return true; }
    return false;
}

// Whether the quad kernel's 32-bit buffer offsets reach every state array of
// a plan of this shape: 4 arrays of level 0, 8 of every coarser level (the
// bound checked at plan creation).
bool quad_offsets_fit(const ginf_bp_desc *d) {
    const int L = d->num_labels;
    const int64_t planes = (L + 3) / 4;
    int64_t w = d->width, h = d->height;
    for (int t = 0; t < d->num_levels; ++t) {
        if (t > 0) { w = (w + 1) / 2; h = (h + 1) / 2; }
        const int64_t whp = round_up((w + 1) / 2, 16);
        const int64_t arr = round_up(planes * round_up(whp * h, ginf::kTileSlots) * 4 * 4, 256);
        if ((t == 0 ? 4 : 8) * arr >= ((int64_t)1 << 32)) return false;
    }
    return true;
}

bool pick_general_shape(int L, int lanes, ginf::SweepShape &sh) {
    if (lanes > 1) return false;
    for (int c : {16, 32, 64, 128})
        if (L <= c) { sh = {c, 1, true, false}; return true; }
    // one wave per node (k_sweep_general_wave), L planes
    if (L <= 1024) { sh = {L, 1, true, false}; return true; }
    return false;
}

}  // namespace

struct ginf_bp_plan_s {
    ginf_bp_desc desc{};
    ginf::SweepShape shape{};
    int levels = 0;
    std::vector<ginf::LevelGeom> geom;
    // per level: data[c], msgs[c][d]
    struct Level { void *data[2]; void *msgs[2][4]; };
    std::vector<Level> lv;
    void *pool = nullptr;
    size_t pool_bytes = 0;
    std::vector<size_t> arr_bytes;  // per level: bytes between consecutive state arrays
    void *table_dev = nullptr;
    std::vector<unsigned char> table_host;
    float s_f = 0, tau_f = 0;
    int s_i = 0, tau_i = 0;
    int device = 0;
    // 16-bit data costs (quad kernel): on by default; a run keeps them when
    // the relayout and pyramid find every cost of every level to be an integer
    // in [0, 65535] (flag_dev[0] stays 0), else it redoes the level data in T.
    // 8-bit data costs at levels 0 and 1 (compact8, on by default): the run
    // first stores those two levels in bytes (16 bits above) and keeps them
    // when every cost there is an integer in [0, 255] (flag_dev[1] stays 0),
    // else it redoes all levels in 16 bits.
    bool compact = true;
    bool compact8 = true;
    bool compact_last = false;
    bool compact8_last = false;
    int prepared_start = -1;  // highest level whose data the last plan_prepare built
    int *flag_dev = nullptr;  // [0]: a cost without a 16-bit form; [1]: one at levels 0-1 without an 8-bit form
    // fused pairs of half-sweeps (ginf_bp_plan_set_pairs) and, per level, the
    // second set of colour-0 outbox arrays they alternate with
    int pair_mode = -1;  // -1: fused where measured faster (label capacity <= 32), 0: off, 1: on
    void *alt_pool = nullptr;
    std::vector<void *> alt;
    // sweep timing (profiling hook)
    bool timing = false;
    struct Timed { hipEvent_t beg, end; double updates; int level; };
    std::vector<Timed> timed;
    std::vector<hipEvent_t> spare;
};

namespace {

size_t elem_size(int dtype) { (void)dtype; return 4; }

// ginf_hbp_host keeps the plan and device buffers of its last problem (one
// entry per process: descriptor, smoothness table and device), so repeated
// synchronous calls on one shape neither reallocate nor re-zero the level
// state; a call with anything different replaces the entry.  The mutex
// serialises callers.  The entry is never freed by a static destructor (the
// HIP runtime may already be gone at exit); ginf_hbp_host_release frees it.
struct HostCache {
    std::mutex mu;
    ginf_bp_plan plan = nullptr;
    void *dd = nullptr, *dl = nullptr;
    ginf_bp_desc desc{};
    int device = -1;
    std::vector<unsigned char> table;
    // upload pipeline: two pinned staging buffers, their events and a stream
    static constexpr size_t kStage = (size_t)32 << 20;
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    hipStream_t up = nullptr;
};

HostCache &host_cache() {
    static HostCache *c = new HostCache;
    return *c;
}

// Frees the problem's device state and, with `all`, the upload pipeline too.
void host_cache_release(HostCache &hc, bool all = true) {
    if (hc.plan) ginf_bp_plan_destroy(hc.plan);
    if (hc.dd) (void)hipFree(hc.dd);
    if (hc.dl) (void)hipFree(hc.dl);
    hc.plan = nullptr;
    hc.dd = hc.dl = nullptr;
    hc.device = -1;
    hc.table.clear();
    if (!all) return;
    for (int k = 0; k < 2; ++k) {
        if (hc.stage[k]) (void)hipHostFree(hc.stage[k]);
        if (hc.stage_ev[k]) (void)hipEventDestroy(hc.stage_ev[k]);
        hc.stage[k] = nullptr;
        hc.stage_ev[k] = nullptr;
    }
    if (hc.up) (void)hipStreamDestroy(hc.up);
    hc.up = nullptr;
}

// Copy n bytes of pageable host memory to the device through the cache's two
// pinned staging buffers: host threads fill one buffer while the DMA engine
// drains the other, so the copy runs at the PCIe rate of pinned memory
// (1.06 GB: 19.6 ms at 54 GB/s against 38.7 ms for hipMemcpy from pageable
// memory on the same box).  Returns after the last byte has arrived.
hipError_t upload_pageable(HostCache &hc, void *dst_dev, const void *src_host, size_t n) {
    hipError_t e = hipSuccess;
    if (!hc.up) e = hipStreamCreateWithFlags(&hc.up, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        if (!hc.stage[k]) e = hipHostMalloc(&hc.stage[k], HostCache::kStage, hipHostMallocDefault);
        if (e == hipSuccess && !hc.stage_ev[k]) e = hipEventCreateWithFlags(&hc.stage_ev[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) return e;
    const unsigned hw = std::thread::hardware_concurrency();
    const int nth = hw >= 4 ? 4 : 1;
    bool used[2] = {false, false};
    size_t k = 0;
    for (size_t off = 0; off < n; off += HostCache::kStage, ++k) {
        const int b = (int)(k & 1);
        const size_t len = n - off < HostCache::kStage ? n - off : HostCache::kStage;
        if (used[b] && (e = hipEventSynchronize(hc.stage_ev[b])) != hipSuccess) return e;
        char *dst = (char *)hc.stage[b];
        const char *src = (const char *)src_host + off;
        if (nth == 1 || len < ((size_t)1 << 20)) {
            std::memcpy(dst, src, len);
        } else {
            // slices 1..nth-1 on helper threads, slice 0 here; a slice whose
            // thread cannot be started is copied here too (no exception may
            // cross the C ABI)
            const size_t per = (len / nth + 4095) & ~(size_t)4095;
            std::vector<std::thread> th;
            for (int t = 1; t < nth && (size_t)t * per < len; ++t) {
                const size_t o = (size_t)t * per, l = len - o < per ? len - o : per;
                try {
                    th.emplace_back([=] { std::memcpy(dst + o, src + o, l); });
                } catch (...) {
                    std::memcpy(dst + o, src + o, l);
                }
            }
            std::memcpy(dst, src, per < len ? per : len);
            for (auto &x : th) x.join();
        }
        if ((e = hipMemcpyAsync((char *)dst_dev + off, dst, len, hipMemcpyHostToDevice, hc.up)) != hipSuccess) return e;
        if ((e = hipEventRecord(hc.stage_ev[b], hc.up)) != hipSuccess) return e;
        used[b] = true;
    }
    return hipStreamSynchronize(hc.up);
}

int validate_desc(const ginf_bp_desc *d) {
    if (!d) return set_error(GINF_ERR_INVALID, "desc is NULL");
    if (d->width < 1 || d->height < 1) return set_error(GINF_ERR_INVALID, "grid %dx%d must be at least 1x1", d->width, d->height);
    if (d->num_labels < 1) return set_error(GINF_ERR_INVALID, "num_labels %d < 1", d->num_labels);
    if (d->dtype != GINF_F32 && d->dtype != GINF_I32) return set_error(GINF_ERR_INVALID, "dtype %d", d->dtype);
    if (d->sm_model < GINF_SMOOTH_EXPLICIT || d->sm_model > GINF_SMOOTH_TRUNC_QUAD)
        return set_error(GINF_ERR_INVALID, "sm_model %d", d->sm_model);
    if (d->num_levels < 1) return set_error(GINF_ERR_INVALID, "num_levels %d < 1", d->num_levels);
    if (d->iters_per_level < 0) return set_error(GINF_ERR_INVALID, "iters_per_level %d < 0", d->iters_per_level);
    if (d->norm != GINF_NORM_MEAN && d->norm != GINF_NORM_MIN) return set_error(GINF_ERR_INVALID, "norm %d", d->norm);
    if (d->kernel < GINF_KERNEL_AUTO || d->kernel > GINF_KERNEL_QUAD) return set_error(GINF_ERR_INVALID, "kernel %d", d->kernel);
    if (d->lanes_per_node < 0 || d->lanes_per_node > 64 || (d->lanes_per_node & (d->lanes_per_node - 1)))
        return set_error(GINF_ERR_INVALID, "lanes_per_node %d (0 or a power of two up to 64)", d->lanes_per_node);
    for (int r : d->reserved)
        if (r != 0) return set_error(GINF_ERR_INVALID, "reserved fields must be zero");
    if ((int64_t)d->width * d->height > (int64_t)1 << 30)
        return set_error(GINF_ERR_UNSUPPORTED, "grid of %lld nodes is too large", (long long)d->width * d->height);
    return GINF_OK;
}

bool is_linear(int model) { return model == GINF_SMOOTH_TRUNC_LINEAR || model == GINF_SMOOTH_POTTS; }

// Interior nodes (1 <= x <= w-2, 1 <= y <= h-2) with (x + y) & 1 == color.
int64_t interior_of_color(int w, int h, int color) {
    if (w < 3 || h < 3) return 0;
    int64_t n = 0;
    for (int y = 1; y <= h - 2; ++y) {
        const int x0 = ((y + color) & 1) ? 1 : 2;  // first x >= 1 with (x + y) & 1 == color
        if (x0 <= w - 2) n += (w - 2 - x0) / 2 + 1;
    }
    return n;
}

// The storage form of the last run's data costs (waits for the device): 0 T,
// 1 16-bit, 2 8-bit at levels 0 and 1 and 16-bit above.
hipError_t data_form_used(ginf_bp_plan p, int *used) {
    *used = 0;
    if (!p->compact_last || !p->flag_dev) return hipSuccess;
    int f[2] = {1, 1};
    hipError_t e = hipMemcpy(f, p->flag_dev, sizeof f, hipMemcpyDeviceToHost);
    if (e == hipSuccess) *used = f[0] != 0 ? 0 : (p->compact8_last && f[1] == 0) ? 2 : 1;
    return e;
}

hipEvent_t take_event(ginf_bp_plan p) {
    if (!p->spare.empty()) { hipEvent_t e = p->spare.back(); p->spare.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

constexpr int kPairMinTiles = 256;

// Whether runs of the plan fuse pairs of half-sweeps (sweep_pair.inc): on
// request, or by default for the label capacities where fused pairs measured
// faster than the unfused half-sweeps (16 and 32).
bool pairs_on(ginf_bp_plan p) {
    if (!p->shape.quad || !ginf::pair_capacity(p->shape.C)) return false;
    return p->pair_mode == 1 || (p->pair_mode == -1 && p->shape.C <= 32);
}

// The second set of colour-0 outbox arrays fused pairs alternate with, per
// level; zeroed once, since border outboxes are never written (as in the
// level's own arrays).
int plan_alloc_pairs(ginf_bp_plan p) {
    if (!pairs_on(p) || p->alt_pool) return GINF_OK;
    size_t total = 0;
    for (int t = 0; t < p->levels; ++t) total += 4 * p->arr_bytes[t];
    GINF_HIP_TRY(hipMalloc(&p->alt_pool, total));
    const hipError_t e = hipMemset(p->alt_pool, 0, total);
    if (e != hipSuccess) {
        (void)hipFree(p->alt_pool);
        p->alt_pool = nullptr;
        return set_error(GINF_ERR_HIP, "pair buffers: %s", hipGetErrorString(e));
    }
    p->alt.clear();
    size_t off = 0;
    for (int t = 0; t < p->levels; ++t) {
        p->alt.push_back((char *)p->alt_pool + off);
        off += 4 * p->arr_bytes[t];
    }
    return GINF_OK;
}

// A prediction that the first half-sweep of the seeded level reads in place
// of seeded outboxes (ginf_bp_plan_run_predicted).
struct PredSource {
    const float *pred;
    int pw, ph, shift;
    float inv_scale;
};

// Level data of levels 0..start from the caller's costs: level 0 (and level
// 1) in one pass over the caller's array, then the coarser levels.  With
// 16-bit storage every value is checked on the device; if any has no exact
// 16-bit form, conditional launches redo the level data in T and the sweeps
// (which read the same flag) take it from there.  Both forms hold the same
// values, and no host synchronisation is needed.  With mp (label-quad layout
// only) the first pass also writes the costs min-pooled over 4 x 4 blocks,
// times mp_scale, as BF16 NHWC (the warm-start network's input); *mp_done
// tells whether it did.
int plan_prepare(ginf_bp_plan p, const void *data_dev, hipStream_t st, int start, uint16_t *mp = nullptr,
                 float mp_scale = 0.f, bool *mp_done = nullptr) {
    using namespace ginf;
    const ginf_bp_desc &d = p->desc;
    const StateLayout ly = state_layout(p->shape, d.num_labels);
    const int dt = d.dtype, C = ly.C, G = ly.G;
    const int W = d.width, H = d.height, L = d.num_labels;
    const bool two = start >= 1;  // a run that starts at level 0 needs no level-1 sums
    const bool with_mp = mp && ly.interleave && ly.G == 4;
    if (mp_done) *mp_done = with_mp;
    // bits: 8 (levels 0 and 1 in bytes, misfits to flag8; levels 2+ in 16
    // bits, misfits to flag), 16 or 32 (T)
    auto prepare = [&](int bits, int *flag, int *flag8, const int *only_if) -> int {
        const bool d16 = bits <= 16, d8 = bits == 8;
        int *f01 = d8 ? flag8 : flag;
        if (with_mp && !only_if)
            GINF_HIP_TRY(launch_relayout_quad_mp(dt, data_dev, W, H, L, ly, p->geom[0], p->lv[0].data[0],
                                                 p->lv[0].data[1], two, p->geom[two ? 1 : 0],
                                                 p->lv[two ? 1 : 0].data[0], p->lv[two ? 1 : 0].data[1], st, d16,
                                                 f01, mp, mp_scale, d8));
        else
            GINF_HIP_TRY(launch_relayout_pyr(dt, data_dev, W, H, L, ly, p->geom[0], p->lv[0].data[0],
                                             p->lv[0].data[1], two, p->geom[two ? 1 : 0], p->lv[two ? 1 : 0].data[0],
                                             p->lv[two ? 1 : 0].data[1], st, d16, f01, only_if, d8));
        for (int t = 2; t <= start; ++t)
            GINF_HIP_TRY(launch_pyramid(dt, p->lv[t - 1].data[0], p->lv[t - 1].data[1], p->geom[t - 1],
                                        p->lv[t].data[0], p->lv[t].data[1], p->geom[t], C, G, st, d16, flag,
                                        only_if, d8 && t == 2));
        return GINF_OK;
    };
    const bool d16 = p->compact && p->shape.quad && p->flag_dev;
    const bool d8 = d16 && p->compact8;
    if (d16) {
        GINF_HIP_TRY(launch_fill_u32((uint32_t *)p->flag_dev, 2, 0u, st));
        if (d8) {
            if (int rc = prepare(8, p->flag_dev, p->flag_dev + 1, nullptr)) return rc;
            if (int rc = prepare(16, p->flag_dev, nullptr, p->flag_dev + 1)) return rc;  // levels 0-1 did not fit
        } else if (int rc = prepare(16, p->flag_dev, nullptr, nullptr)) {
            return rc;
        }
        if (int rc = prepare(32, nullptr, nullptr, p->flag_dev)) return rc;
    } else if (int rc = prepare(32, nullptr, nullptr, nullptr)) {
        return rc;
    }
    p->compact_last = d16;
    p->compact8_last = d8;
    p->prepared_start = start;
    return GINF_OK;
}

// The HBP driver.  iters[t] half-sweeps run at level t for t <= start;
// levels above `start` are skipped.  When `seeded`, level `start` begins from
// the outboxes already in its arrays (written by ginf_bp_plan_seed) instead
// of zeros.  ginf_bp_plan_run is start = levels-1, iters[t] = the
// descriptor's count, not seeded: decodeHbp (src/bp.cc:86-180).  With
// `prepared` the level data of the last plan_prepare is used as it is.
int plan_run_ex(ginf_bp_plan p, const void *data_dev, int32_t *labels_dev, hipStream_t st, int start,
                const int *iters, bool seeded, const PredSource *pred = nullptr, bool prepared = false) {
    using namespace ginf;
    const ginf_bp_desc &d = p->desc;
    const StateLayout ly = state_layout(p->shape, d.num_labels);
    const int dt = d.dtype, C = ly.C, G = ly.G;
    const int W = d.width, H = d.height, L = d.num_labels;
    if (!prepared) {
        if (int rc = plan_prepare(p, data_dev, st, start)) return rc;
    }
    const bool d16 = p->compact_last;
    for (int t = start; t >= 0; --t) {
        const LevelGeom &g = p->geom[t];
        const int64_t per = (int64_t)C * g.nplane * G;
        const int N = iters[t];
        // With two or more half-sweeps both colours move, so the first
        // half-sweep can read the level above directly (fused upsample) or,
        // at an unseeded top level, zero messages without reading them.
        const bool top = t == start;
        const bool fused = N >= 2 && !(top && seeded);
        if (!fused) {
            if (top) {
                if (!seeded)
                    for (int c = 0; c < 2; ++c)
                        for (int k = 0; k < 4; ++k) GINF_HIP_TRY(launch_fill_u32(p->lv[t].msgs[c][k], per, 0u, st));
            } else {
                UpsampleArgs u{};
                for (int c = 0; c < 2; ++c)
                    for (int k = 0; k < 4; ++k) { u.coarse[c][k] = p->lv[t + 1].msgs[c][k]; u.fine[c][k] = p->lv[t].msgs[c][k]; }
                u.cg = p->geom[t + 1];
                u.fg = g;
                u.C = C;
                u.G = G;
                GINF_HIP_TRY(launch_upsample(dt, u, st));
            }
        }
        if (t == 0 && N < 2) {
            // Interior nodes never updated at level 0 keep a zero belief: label 0
            // (src/bp.cc:168-173).
            GINF_HIP_TRY(launch_zero_labels(labels_dev, W, H, 0, st));
            if (N == 0) GINF_HIP_TRY(launch_zero_labels(labels_dev, W, H, 1, st));
        }
        auto make_args = [&](int i) {
            SweepArgs a{};
            const int color = (i + 1) & 1;  // (x + y + i) odd (src/bp.cc:150)
            a.data = p->lv[t].data[color];
            for (int k = 0; k < 4; ++k) a.out[k] = p->lv[t].msgs[color][k];
            const int src_level = (fused && i == 0 && !top) ? t + 1 : t;
            for (int c = 0; c < 2; ++c)
                for (int k = 0; k < 4; ++k) a.in[c][k] = p->lv[src_level].msgs[c][k];
            a.src = src_level == t ? kSrcSame : kSrcCoarse;
            a.stride = a.in_stride = (int64_t)ginf::kTileSlots * G;
            a.tile = (int64_t)C * ginf::kTileSlots * G;
            a.w = g.w; a.h = g.h; a.whp = g.whp; a.nslots = g.nslots;
            a.in_whp = p->geom[src_level].whp;
            a.color = color;
            a.num_labels = L;
            a.norm = d.norm;
            a.labels = labels_dev;
            a.label_w = W;
            a.s_f = p->s_f; a.tau_f = p->tau_f; a.s_i = p->s_i; a.tau_i = p->tau_i;
            a.table = p->table_dev;
            // byte view for the quad kernel
            a.arr_bytes = (uint32_t)p->arr_bytes[t];
            a.data_bytes = (uint32_t)p->arr_bytes[t];
            a.plane_bytes = (uint32_t)(ginf::kTileSlots * G * 4);
            a.tile_bytes = (uint32_t)(a.tile * 4);
            a.data_plane_bytes = a.plane_bytes / 2;
            a.fit_flag = d16 ? p->flag_dev : nullptr;
            a.fit8_flag = (d16 && p->compact8_last && t <= 1) ? p->flag_dev + 1 : nullptr;
            a.out_base = p->lv[t].msgs[color][0];
            if (src_level == t) {
                a.in_base = p->lv[t].msgs[1 - color][0];
                a.in_arr_bytes = a.arr_bytes;
                a.in_plane_bytes = a.plane_bytes;
                a.in_span = (uint32_t)(4 * p->arr_bytes[t]);
            } else {
                a.in_base = p->lv[src_level].msgs[0][0];
                a.in_arr_bytes = (uint32_t)p->arr_bytes[src_level];
                a.in_plane_bytes = a.plane_bytes;
                a.in_span = (uint32_t)(8 * p->arr_bytes[src_level]);
            }
            return a;
        };
        // Fused pairs of half-sweeps (quad layout): colour A's messages of
        // half-sweep i reach colour B at i + 1 through LDS.  A pair's stage A
        // reads colour B's outboxes of the previous pair at its tile edges
        // while neighbouring tiles write colour B's new outboxes, so
        // successive pairs alternate colour B's outboxes (always colour 0)
        // between the level's arrays and a second set, arranged so that the
        // level's last pair writes the level's own arrays.  A pair never
        // reads and writes the same set: where the first pair would read the
        // level's own colour-B arrays (a seeded top level) and an odd number
        // of pairs would make it write them too, the first two half-sweeps
        // run unfused.
        // A level with fewer tiles than compute units runs unfused: two short
        // launches are quicker there than one fused launch whose workgroups
        // each wait for two HBM round trips (128x96x16 plain BP, 10
        // half-sweeps: 0.044 ms unfused, 0.051 ms fused, 48 tiles).
        const bool pairs = pairs_on(p) && p->alt_pool && g.w >= 3 && g.h >= 3 &&
                           ginf::pair_tiles_f32(p->shape.C, g.whp, g.h) >= kPairMinTiles;
        // With a prediction source (top level, N >= 2) half-sweep 0 reads the
        // prediction and half-sweeps 0 and 1 run unfused; the pairs after them
        // read the level's own colour-B arrays first, so their count must be
        // even as for a seeded top level.
        const bool from_pred = top && pred != nullptr;
        int npairs = pairs ? N / 2 : 0, first_pair = 0;
        if (from_pred && npairs > 0) {
            first_pair = 2;
            --npairs;
            if (npairs % 2 == 1) { first_pair = 4; --npairs; }
        } else if (npairs % 2 == 1 && top && seeded) {
            first_pair = 2;
            --npairs;
        }
        int k = 0;
        bool prev_alt = false;
        for (int i = 0; i < N;) {
            const bool zeroin = fused && i == 0 && top;
            const bool pair = pairs && i >= first_pair && i + 1 < N;
            const bool labels_i = (t == 0) && (i >= N - 2);
            const bool labels_j = pair && (t == 0) && (i + 1 >= N - 2);
            const SweepArgs a = make_args(i);
            ginf_bp_plan_s::Timed tm{nullptr, nullptr, 0.0, t};
            // Only the steady-state kernel instance (no zero-input, no labels) is
            // timed, so the figure matches a per-kernel profiler summary.
            if (p->timing && !zeroin && !labels_i && !labels_j && !(from_pred && i == 0)) {
                tm.beg = take_event(p);
                tm.end = take_event(p);
                tm.updates = (double)interior_of_color(g.w, g.h, a.color) * L;
                if (pair) tm.updates += (double)interior_of_color(g.w, g.h, 1 - a.color) * L;
                if (tm.beg) GINF_HIP_TRY(hipEventRecord(tm.beg, st));
            }
            if (from_pred && i == 0) {
                SweepArgs ap = a;
                ap.src = kSrcSame;
                ap.in_base = pred->pred;
                ap.in_span = (uint32_t)((size_t)pred->pw * pred->ph * 4 * L * sizeof(float));
                ap.in_plane_bytes = 16;  // label quad j of a prediction row
                ap.pred_w = pred->pw;
                ap.pred_h = pred->ph;
                ap.pred_shift = pred->shift;
                ap.pred_inv_scale = pred->inv_scale;
                GINF_HIP_TRY(launch_sweep_seeded(p->shape, ap, labels_i, d16, st));
            } else if (pair) {
                SweepArgs aa = a;
                SweepArgs b = make_args(i + 1);
                if (k > 0 && prev_alt) aa.in_base = p->alt[t];  // the previous pair's colour-B outboxes
                const bool dst_alt = (npairs - 1 - k) % 2 == 1;
                if (dst_alt) b.out_base = p->alt[t];
                prev_alt = dst_alt;
                ++k;
                const int flags = (labels_i ? 1 : 0) | (labels_j ? 2 : 0) | (i + 2 >= N ? 4 : 0);
                GINF_HIP_TRY(dt == GINF_F32 ? launch_sweep_pair_f32(p->shape.C, aa, b, zeroin, flags, st)
                                            : launch_sweep_pair_i32(p->shape.C, aa, b, zeroin, flags, st));
            } else {
                GINF_HIP_TRY(launch_sweep(dt, p->shape, a, zeroin, labels_i, st, d16));
            }
            if (tm.beg && tm.end) {
                GINF_HIP_TRY(hipEventRecord(tm.end, st));
                p->timed.push_back(tm);
            }
            i += pair ? 2 : 1;
        }
    }
    return GINF_OK;
}


int plan_run(ginf_bp_plan p, const void *data_dev, int32_t *labels_dev, hipStream_t st) {
    std::vector<int> iters((size_t)p->levels, p->desc.iters_per_level);
    return plan_run_ex(p, data_dev, labels_dev, st, p->levels - 1, iters.data(), false);
}
}  // namespace

extern "C" {

const char *ginf_last_error(void) { return g_last_error.c_str(); }

const char *ginf_version(void) { return "gridinfcuda_amd 0.1 (gfx950)"; }

int ginf_bp_plan_create(const ginf_bp_desc *desc, const void *sm_table_host, ginf_bp_plan *out) {
    if (!out) return set_error(GINF_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate_desc(desc);
    if (rc) return rc;
    if (!sm_table_host) return set_error(GINF_ERR_INVALID, "sm_table_host is NULL");
    const int L = desc->num_labels;
    ginf::SweepShape sh{};
    bool ok;
    if (!is_linear(desc->sm_model)) {
        ok = desc->kernel != GINF_KERNEL_QUAD && pick_general_shape(L, desc->lanes_per_node, sh);
    } else if (desc->kernel == GINF_KERNEL_QUAD) {
        ok = pick_quad_shape(L, sh);
    } else if (desc->kernel == GINF_KERNEL_CHUNKED || desc->lanes_per_node != 0) {
        ok = pick_linear_shape(L, desc->lanes_per_node, sh);
    } else if ((desc->dtype == GINF_F32 && desc->norm == GINF_NORM_MEAN && L <= 256) ||
               (L <= 32 && quad_offsets_fit(desc))) {
        // FP32 mean normalisation rounds differently under any other order:
        // keep the reference's sequential per-direction order.  Up to 32
        // labels the quad kernel (with fused pairs) is also the faster one for
        // integer arithmetic (1920x1080x16 HBP: 1.29-1.31 ms against 1.65-1.68
        // ms chunked; x32: 2.68-2.94 against 3.10-3.12; 128x96x16: 0.044-0.046
        // against 0.060).
        ok = pick_quad_shape(L, sh);
    } else {
        ok = pick_linear_shape(L, 0, sh);
    }
    if (!ok)
        return set_error(GINF_ERR_UNSUPPORTED, "no kernel for L=%d, model %d, lanes_per_node %d", L, desc->sm_model,
                         desc->lanes_per_node);
    ginf_bp_plan p = new (std::nothrow) ginf_bp_plan_s;
    if (!p) return set_error(GINF_ERR_NOMEM, "host allocation failed");
    p->desc = *desc;
    p->shape = sh;
    p->levels = desc->num_levels;
    hipGetDevice(&p->device);
    const size_t es = elem_size(desc->dtype);
    p->table_host.resize((size_t)L * L * es);
    memcpy(p->table_host.data(), sm_table_host, p->table_host.size());
    // O(L) parameters s = V(0,1) and tau = V(0,L-1) (src/bp.cc:35,43); with
    // L == 1 the scans are empty and s is unused.
    if (desc->dtype == GINF_F32) {
        const float *tb = (const float *)sm_table_host;
        p->s_f = L > 1 ? tb[L] : 0.f;
        p->tau_f = tb[(size_t)L * (L - 1)];
    } else {
        const int *tb = (const int *)sm_table_host;
        p->s_i = L > 1 ? tb[L] : 0;
        p->tau_i = tb[(size_t)L * (L - 1)];
    }
    // Level geometry: ceil halving (src/bp.cc:101-102).
    int w = desc->width, h = desc->height;
    size_t total = 0;
    std::vector<size_t> offs;
    for (int t = 0; t < p->levels; ++t) {
        if (t > 0) { w = (w + 1) / 2; h = (h + 1) / 2; }
        ginf::LevelGeom g{};
        g.w = w; g.h = h;
        g.whp = (int)round_up((w + 1) / 2, 16);
        g.nslots = g.whp * h;
        g.nplane = round_up(g.nslots, ginf::kTileSlots);
        p->geom.push_back(g);
        const ginf::StateLayout ly = ginf::state_layout(sh, L);
        const size_t arr = (size_t)ly.C * g.nplane * ly.G * es;
        const size_t stride = (GINF_ARRAY_STAGGER > 0 && arr >= ((size_t)1 << 20))
                                  ? round_up(arr, (size_t)1 << 20) + GINF_ARRAY_STAGGER
                                  : round_up(arr, 256);
        for (int k = 0; k < 10; ++k) { offs.push_back(total); total += stride; }
        p->arr_bytes.push_back(stride);
    }
    total += round_up((size_t)L * L * es, 256);
    if (sh.quad) {
        // The quad kernel addresses a level's outbox arrays with 32-bit buffer
        // offsets: 4 arrays of the level itself, 8 of the level above.
        bool fits = 4 * (uint64_t)p->arr_bytes[0] < ((uint64_t)1 << 32);
        for (int t = 1; t < p->levels; ++t) fits = fits && 8 * (uint64_t)p->arr_bytes[t] < ((uint64_t)1 << 32);
        if (!fits) {
            delete p;
            return set_error(GINF_ERR_UNSUPPORTED, "grid too large for the quad kernel's 32-bit offsets "
                                                   "(W*H*L must stay below about 2^29); use GINF_KERNEL_CHUNKED");
        }
    }
    p->pool_bytes = total;
    hipError_t e = hipMalloc(&p->pool, total);
    if (e != hipSuccess) {
        delete p;
        return set_error(GINF_ERR_NOMEM, "hipMalloc(%zu) failed: %s", total, hipGetErrorString(e));
    }
    // Zero everything once: border outboxes are never written and stay 0.
    e = hipMemset(p->pool, 0, total);
    if (e == hipSuccess) {
        char *base = (char *)p->pool;
        p->lv.resize(p->levels);
        size_t k = 0;
        for (int t = 0; t < p->levels; ++t) {
            for (int c = 0; c < 2; ++c) p->lv[t].data[c] = base + offs[k++];
            for (int c = 0; c < 2; ++c)
                for (int dd = 0; dd < 4; ++dd) p->lv[t].msgs[c][dd] = base + offs[k++];
        }
        p->table_dev = base + total - round_up((size_t)L * L * es, 256);
        e = hipMemcpy(p->table_dev, sm_table_host, (size_t)L * L * es, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && p->shape.quad) {
        e = hipMalloc((void **)&p->flag_dev, 2 * sizeof(int));
    }
    if (e != hipSuccess) {
        if (p->flag_dev) hipFree(p->flag_dev);
        hipFree(p->pool);
        delete p;
        return set_error(GINF_ERR_HIP, "plan set-up failed: %s", hipGetErrorString(e));
    }
    if (int rc2 = plan_alloc_pairs(p)) {
        ginf_bp_plan_destroy(p);
        return rc2;
    }
    *out = p;
    return GINF_OK;
}

int ginf_bp_plan_destroy(ginf_bp_plan plan) {
    if (!plan) return GINF_OK;
    for (auto &t : plan->timed) { if (t.beg) hipEventDestroy(t.beg); if (t.end) hipEventDestroy(t.end); }
    for (auto e : plan->spare) hipEventDestroy(e);
    if (plan->pool) hipFree(plan->pool);
    if (plan->flag_dev) hipFree(plan->flag_dev);
    if (plan->alt_pool) hipFree(plan->alt_pool);
    delete plan;
    return GINF_OK;
}

int ginf_bp_plan_workspace_bytes(ginf_bp_plan plan, size_t *bytes) {
    if (!plan || !bytes) return set_error(GINF_ERR_INVALID, "NULL argument");
    *bytes = plan->pool_bytes;
    return GINF_OK;
}

int ginf_bp_plan_set_compact_data(ginf_bp_plan plan, int enable) {
    if (!plan) return set_error(GINF_ERR_INVALID, "NULL plan");
    if (enable < 0 || enable > 2) return set_error(GINF_ERR_INVALID, "compact data mode %d (0, 1 or 2)", enable);
    plan->compact = enable != 0;
    plan->compact8 = enable == 1;
    return GINF_OK;
}

int ginf_bp_plan_set_pairs(ginf_bp_plan plan, int enable) {
    if (!plan) return set_error(GINF_ERR_INVALID, "NULL plan");
    if (enable < -1 || enable > 1) return set_error(GINF_ERR_INVALID, "pairs mode %d (-1, 0 or 1)", enable);
    plan->pair_mode = enable;
    return plan_alloc_pairs(plan);
}

int ginf_bp_plan_pairs(ginf_bp_plan plan, int *enabled) {
    if (!plan || !enabled) return set_error(GINF_ERR_INVALID, "NULL argument");
    *enabled = pairs_on(plan) ? 1 : 0;
    return GINF_OK;
}

int ginf_bp_plan_compact_data(ginf_bp_plan plan, int *last_run_compact) {
    if (!plan || !last_run_compact) return set_error(GINF_ERR_INVALID, "NULL argument");
    GINF_HIP_TRY(data_form_used(plan, last_run_compact));
    return GINF_OK;
}

int ginf_bp_plan_set_sweep_timing(ginf_bp_plan plan, int enable) {
    if (!plan) return set_error(GINF_ERR_INVALID, "NULL plan");
    plan->timing = enable != 0;
    return GINF_OK;
}

int ginf_bp_plan_sweep_stats(ginf_bp_plan p, double *total_ms, int64_t *launches, double *updates, int reset) {
    if (!p || !total_ms || !launches || !updates) return set_error(GINF_ERR_INVALID, "NULL argument");
    double ms = 0.0, up = 0.0;
    for (auto &t : p->timed) {
        GINF_HIP_TRY(hipEventSynchronize(t.end));
        float e = 0.f;
        GINF_HIP_TRY(hipEventElapsedTime(&e, t.beg, t.end));
        ms += e;
        up += t.updates;
    }
    *total_ms = ms;
    *launches = (int64_t)p->timed.size();
    *updates = up;
    if (reset) {
        for (auto &t : p->timed) { p->spare.push_back(t.beg); p->spare.push_back(t.end); }
        p->timed.clear();
    }
    return GINF_OK;
}

int ginf_bp_plan_sweep_level_updates(ginf_bp_plan p, double *updates_by_level) {
    if (!p || !updates_by_level) return set_error(GINF_ERR_INVALID, "NULL argument");
    for (int t = 0; t < p->levels; ++t) updates_by_level[t] = 0.0;
    for (auto &t : p->timed) updates_by_level[t.level] += t.updates;
    return GINF_OK;
}

int ginf_bp_plan_updates_per_run(ginf_bp_plan p, double *updates) {
    if (!p || !updates) return set_error(GINF_ERR_INVALID, "NULL argument");
    double u = 0.0;
    const int N = p->desc.iters_per_level;
    for (int t = 0; t < p->levels; ++t)
        for (int i = 0; i < N; ++i)
            u += (double)interior_of_color(p->geom[t].w, p->geom[t].h, (i + 1) & 1) * p->desc.num_labels;
    *updates = u;
    return GINF_OK;
}

int ginf_bp_plan_level_dims(ginf_bp_plan plan, int level, int *w, int *h) {
    if (!plan || !w || !h) return set_error(GINF_ERR_INVALID, "NULL argument");
    if (level < 0 || level >= plan->levels) return set_error(GINF_ERR_INVALID, "level %d out of range", level);
    *w = plan->geom[level].w;
    *h = plan->geom[level].h;
    return GINF_OK;
}

int ginf_bp_plan_run(ginf_bp_plan plan, const void *data_dev, int32_t *labels_dev, void *stream) {
    if (!plan || !data_dev || !labels_dev) return set_error(GINF_ERR_INVALID, "NULL argument");
    return plan_run(plan, data_dev, labels_dev, (hipStream_t)stream);
}

int ginf_bp_plan_run_schedule(ginf_bp_plan p, const void *data_dev, int32_t *labels_dev, int start_level,
                              const int *iters, int seeded, void *stream) {
    if (!p || !data_dev || !labels_dev || !iters) return set_error(GINF_ERR_INVALID, "NULL argument");
    if (start_level < 0 || start_level >= p->levels) return set_error(GINF_ERR_INVALID, "start level %d", start_level);
    for (int t = 0; t <= start_level; ++t)
        if (iters[t] < 0) return set_error(GINF_ERR_INVALID, "iters[%d] < 0", t);
    return plan_run_ex(p, data_dev, labels_dev, (hipStream_t)stream, start_level, iters, seeded != 0);
}

int ginf_bp_plan_seed(ginf_bp_plan p, int level, const float *pred_dev, int pw, int ph, int shift, float scale,
                      void *stream) {
    if (!p || !pred_dev || pw < 1 || ph < 1 || shift < 0 || shift > 8 || scale == 0.f)
        return set_error(GINF_ERR_INVALID, "bad seed arguments");
    if (level < 0 || level >= p->levels) return set_error(GINF_ERR_INVALID, "level %d out of range", level);
    if (p->desc.dtype != GINF_F32) return set_error(GINF_ERR_UNSUPPORTED, "seeding needs FP32 messages");
    const ginf::StateLayout ly = ginf::state_layout(p->shape, p->desc.num_labels);
    void *m[2][4];
    for (int c = 0; c < 2; ++c)
        for (int d = 0; d < 4; ++d) m[c][d] = p->lv[level].msgs[c][d];
    GINF_HIP_TRY(ginf::launch_inject(pred_dev, m, p->geom[level], ly.C, ly.G, ly.interleave, p->desc.num_labels, shift,
                                     pw, ph, 1.f / scale, (hipStream_t)stream));
    return GINF_OK;
}

int ginf_bp_plan_prepare(ginf_bp_plan p, const void *data_dev, int start_level, uint16_t *minpool_out,
                         float minpool_scale, void *stream) {
    if (!p || !data_dev) return set_error(GINF_ERR_INVALID, "NULL argument");
    if (start_level < 0 || start_level >= p->levels) return set_error(GINF_ERR_INVALID, "start level %d", start_level);
    bool done = false;
    if (int rc = plan_prepare(p, data_dev, (hipStream_t)stream, start_level, minpool_out, minpool_scale, &done))
        return rc;
    if (minpool_out && !done) {
        // layouts other than label quads: the separate min-pool kernel
        const ginf_bp_desc &d = p->desc;
        if (ginf_minpool_costs(data_dev, d.dtype, d.width, d.height, d.num_labels, 4, minpool_scale, minpool_out,
                               stream) != GINF_OK)
            return set_error(GINF_ERR_HIP, "min-pool kernel failed");
    }
    return GINF_OK;
}

int ginf_bp_plan_run_predicted(ginf_bp_plan p, const void *data_dev, int32_t *labels_dev, int start_level,
                               const int *iters, const float *pred_dev, int pw, int ph, int shift, float scale,
                               int flags, void *stream) {
    if (!p || !data_dev || !labels_dev || !iters || !pred_dev) return set_error(GINF_ERR_INVALID, "NULL argument");
    if (start_level < 0 || start_level >= p->levels) return set_error(GINF_ERR_INVALID, "start level %d", start_level);
    for (int t = 0; t <= start_level; ++t)
        if (iters[t] < 0) return set_error(GINF_ERR_INVALID, "iters[%d] < 0", t);
    if (pw < 1 || ph < 1 || shift < 0 || shift > 8 || scale == 0.f)
        return set_error(GINF_ERR_INVALID, "bad prediction arguments");
    if (flags & ~GINF_RUN_PREPARED) return set_error(GINF_ERR_INVALID, "unknown flags 0x%x", flags);
    const bool prepared = (flags & GINF_RUN_PREPARED) != 0;
    if (prepared && p->prepared_start < start_level)
        return set_error(GINF_ERR_INVALID, "level %d was not prepared (ginf_bp_plan_prepare)", start_level);
    if (p->desc.dtype != GINF_F32) return set_error(GINF_ERR_UNSUPPORTED, "seeding needs FP32 messages");
    const int L = p->desc.num_labels;
    const bool direct = p->shape.quad && L % 4 == 0 && iters[start_level] >= 2 && ((uintptr_t)pred_dev & 15) == 0 &&
                        (size_t)pw * ph * 4 * L * sizeof(float) < ((size_t)1 << 32);
    if (!direct) {
        // Other layouts, or fewer than two half-sweeps at the seeded level
        // (where colour B's seeded outboxes stay in use): seed, then run.
        if (int rc = ginf_bp_plan_seed(p, start_level, pred_dev, pw, ph, shift, scale, stream)) return rc;
        return plan_run_ex(p, data_dev, labels_dev, (hipStream_t)stream, start_level, iters, true, nullptr, prepared);
    }
    const PredSource src{pred_dev, pw, ph, shift, 1.f / scale};
    return plan_run_ex(p, data_dev, labels_dev, (hipStream_t)stream, start_level, iters, true, &src, prepared);
}

int ginf_bp_plan_block_mean_msgs(ginf_bp_plan p, int level, int f, float scale, float *out_dev, void *stream) {
    if (!p || !out_dev || f < 1) return set_error(GINF_ERR_INVALID, "bad arguments");
    if (level < 0 || level >= p->levels) return set_error(GINF_ERR_INVALID, "level %d out of range", level);
    if (p->desc.dtype != GINF_F32) return set_error(GINF_ERR_UNSUPPORTED, "FP32 plans only");
    const ginf::StateLayout ly = ginf::state_layout(p->shape, p->desc.num_labels);
    void *m[2][4];
    for (int c = 0; c < 2; ++c)
        for (int d = 0; d < 4; ++d) m[c][d] = p->lv[level].msgs[c][d];
    GINF_HIP_TRY(ginf::launch_block_mean_state(m, p->geom[level], ly.C, ly.G, ly.interleave, p->desc.num_labels, f,
                                               scale, out_dev, (hipStream_t)stream));
    return GINF_OK;
}

int ginf_bp_plan_dump_msgs(ginf_bp_plan p, int level, void *msgs_host) {
    if (!p || !msgs_host) return set_error(GINF_ERR_INVALID, "NULL argument");
    if (level < 0 || level >= p->levels) return set_error(GINF_ERR_INVALID, "level %d out of range", level);
    const ginf::LevelGeom &g = p->geom[level];
    const int L = p->desc.num_labels;
    const ginf::StateLayout ly = ginf::state_layout(p->shape, L);
    const int C = ly.C, G = ly.G;
    const size_t n = (size_t)C * g.nplane * G;
    std::vector<uint32_t> buf(n);
    uint32_t *dst = (uint32_t *)msgs_host;
    GINF_HIP_TRY(hipDeviceSynchronize());
    for (int c = 0; c < 2; ++c)
        for (int d = 0; d < 4; ++d) {
            GINF_HIP_TRY(hipMemcpy(buf.data(), p->lv[level].msgs[c][d], n * 4, hipMemcpyDeviceToHost));
            for (int y = 0; y < g.h; ++y)
                for (int x = 0; x < g.w; ++x) {
                    if (((x + y) & 1) != c) continue;
                    const size_t slot = (size_t)y * g.whp + (x >> 1);
                    for (int l = 0; l < L; ++l) {
                        const int q = ly.interleave ? l % G : l / C;
                        const int i = ly.interleave ? l / G : l - q * C;
                        dst[(size_t)x + (size_t)g.w * ((size_t)y + (size_t)g.h * ((size_t)d + 4u * l))] =
                            buf[ginf::elem_index(slot, i, q, C, G)];
                    }
                }
        }
    return GINF_OK;
}

int ginf_bp_plan_dump_data(ginf_bp_plan p, int level, void *data_host) {
    if (!p || !data_host) return set_error(GINF_ERR_INVALID, "NULL argument");
    if (level < 0 || level >= p->levels) return set_error(GINF_ERR_INVALID, "level %d out of range", level);
    const ginf::LevelGeom &g = p->geom[level];
    const int L = p->desc.num_labels;
    const ginf::StateLayout ly = ginf::state_layout(p->shape, L);
    const int C = ly.C, G = ly.G;
    const size_t n = (size_t)C * g.nplane * G;
    std::vector<uint32_t> buf(n);
    GINF_HIP_TRY(hipDeviceSynchronize());
    int form = 0;
    GINF_HIP_TRY(data_form_used(p, &form));
    const int bytes = form == 0 ? 4 : (form == 2 && level <= 1) ? 1 : 2;
    std::vector<uint16_t> buf16(bytes == 2 ? n : 0);
    std::vector<uint8_t> buf8(bytes == 1 ? n : 0);
    uint32_t *dst = (uint32_t *)data_host;
    for (int c = 0; c < 2; ++c) {
        if (bytes < 4) {  // 16- or 8-bit storage of the last run: widen to T
            if (bytes == 2) GINF_HIP_TRY(hipMemcpy(buf16.data(), p->lv[level].data[c], n * 2, hipMemcpyDeviceToHost));
            else GINF_HIP_TRY(hipMemcpy(buf8.data(), p->lv[level].data[c], n, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < n; ++k) {
                const uint32_t v = bytes == 2 ? buf16[k] : buf8[k];
                if (p->desc.dtype == GINF_F32) {
                    const float f = (float)v;
                    std::memcpy(&buf[k], &f, 4);
                } else {
                    buf[k] = v;
                }
            }
        } else {
            GINF_HIP_TRY(hipMemcpy(buf.data(), p->lv[level].data[c], n * 4, hipMemcpyDeviceToHost));
        }
        for (int y = 0; y < g.h; ++y)
            for (int x = 0; x < g.w; ++x) {
                if (((x + y) & 1) != c) continue;
                const size_t slot = (size_t)y * g.whp + (x >> 1);
                for (int l = 0; l < L; ++l) {
                    const int q = ly.interleave ? l % G : l / C;
                    const int i = ly.interleave ? l / G : l - q * C;
                    dst[(size_t)x + (size_t)g.w * ((size_t)y + (size_t)g.h * (size_t)l)] =
                        buf[ginf::elem_index(slot, i, q, C, G)];
                }
            }
    }
    return GINF_OK;
}

int ginf_hbp_host(const ginf_bp_desc *desc, const void *data_host, const void *sm_table_host, int32_t *labels_host) {
    int rc = validate_desc(desc);
    if (rc) return rc;
    if (!data_host || !labels_host || !sm_table_host) return set_error(GINF_ERR_INVALID, "NULL argument");
    int dev = 0;
    GINF_HIP_TRY(hipGetDevice(&dev));
    const size_t nd = (size_t)desc->width * desc->height * desc->num_labels * 4;
    const size_t nl = (size_t)desc->width * desc->height * 4;
    const size_t nt = (size_t)desc->num_labels * desc->num_labels * elem_size(desc->dtype);
    HostCache &hc = host_cache();
    std::lock_guard<std::mutex> lock(hc.mu);
    const bool hit = hc.plan && hc.device == dev && std::memcmp(&hc.desc, desc, sizeof(ginf_bp_desc)) == 0 &&
                     hc.table.size() == nt && std::memcmp(hc.table.data(), sm_table_host, nt) == 0;
    if (!hit) {
        // a new problem; the upload pipeline is kept unless the device changed
        host_cache_release(hc, hc.device != dev);
        rc = ginf_bp_plan_create(desc, sm_table_host, &hc.plan);
        if (rc) return rc;
        hipError_t e = hipMalloc(&hc.dd, nd);
        if (e == hipSuccess) e = hipMalloc(&hc.dl, nl);
        if (e != hipSuccess) {
            host_cache_release(hc);
            return set_error(GINF_ERR_HIP, "ginf_hbp_host: %s", hipGetErrorString(e));
        }
        hc.desc = *desc;
        hc.device = dev;
        hc.table.assign((const unsigned char *)sm_table_host, (const unsigned char *)sm_table_host + nt);
    }
    hipError_t e = upload_pageable(hc, hc.dd, data_host, nd);
    if (e == hipSuccess) e = hipMemcpy(hc.dl, labels_host, nl, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = plan_run(hc.plan, hc.dd, (int32_t *)hc.dl, nullptr);
        if (rc == GINF_OK) e = hipMemcpy(labels_host, hc.dl, nl, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess || rc != GINF_OK) {
        host_cache_release(hc);
        if (e != hipSuccess) return set_error(GINF_ERR_HIP, "ginf_hbp_host: %s", hipGetErrorString(e));
    }
    return rc;
}

int ginf_hbp_host_release(void) {
    HostCache &hc = host_cache();
    std::lock_guard<std::mutex> lock(hc.mu);
    host_cache_release(hc);
    return GINF_OK;
}

int ginf_energy_f64(int width, int height, int num_labels, int dtype, const void *data_dev, const void *table_dev,
                    const int32_t *labels_dev, double *energy_host, void *stream) {
    if (!data_dev || !table_dev || !labels_dev || !energy_host) return set_error(GINF_ERR_INVALID, "NULL argument");
    if (width < 1 || height < 1 || num_labels < 1) return set_error(GINF_ERR_INVALID, "bad sizes");
    if (dtype != GINF_F32 && dtype != GINF_I32) return set_error(GINF_ERR_INVALID, "dtype %d", dtype);
    if (width < 3 || height < 3) { *energy_host = 0.0; return GINF_OK; }
    const int nblocks = 1024;
    double *buf = nullptr;
    GINF_HIP_TRY(hipMalloc(&buf, sizeof(double) * (nblocks + 1)));
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = ginf::launch_energy(dtype, data_dev, table_dev, labels_dev, width, height, num_labels, buf, nblocks,
                                       buf + nblocks, st);
    if (e == hipSuccess) e = hipMemcpyAsync(energy_host, buf + nblocks, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(buf);
    if (e != hipSuccess) return set_error(GINF_ERR_HIP, "energy: %s", hipGetErrorString(e));
    if (*energy_host != *energy_host) return set_error(GINF_ERR_INVALID, "a label is outside [0, %d)", num_labels);
    return GINF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The energy without allocation (no reference counterpart): the per-block
// partial sums and the total of k_energy_partial / k_energy_final live in a
// fixed device array instead of a hipMalloc'd buffer, for callers that cannot
// allocate (tight serving loops such as the warm start's per-instance
// fallback decisions, code that must not synchronise the device through
// hipMalloc/hipFree).  The array's size does not depend on the grid: the
// partial-sum kernel is a grid-stride loop over kEnergyBlocks blocks.
namespace {

constexpr int kEnergyBlocks = 1024;
constexpr int kEnergyScratch = kEnergyBlocks + 1;  // partials, then the total (8,200 bytes)
static __device__ __align__(256) double scratch[kEnergyScratch];

// Device address of `scratch` on each device, looked up on the first call
// made with that device current (each GPU holds its own copy of the array).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

int energy_scratch(int dev, double **out) {
    if (dev < 0 || dev >= kMaxDevices) return set_error(GINF_ERR_UNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(GINF_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_dev[dev] = (double *)ptr;
    }
    *out = g_scratch_dev[dev];
    return GINF_OK;
}

}  // namespace

extern "C" int ginf_energy_f64_noalloc(int width, int height, int num_labels, int dtype, const void *data_dev,
                                       const void *table_dev, const int32_t *labels_dev, double *energy_host,
                                       void *stream) {
    if (!data_dev || !table_dev || !labels_dev || !energy_host) return set_error(GINF_ERR_INVALID, "NULL argument");
    if (width < 1 || height < 1 || num_labels < 1) return set_error(GINF_ERR_INVALID, "bad sizes");
    if (dtype != GINF_F32 && dtype != GINF_I32) return set_error(GINF_ERR_INVALID, "dtype %d", dtype);
    if (width < 3 || height < 3) { *energy_host = 0.0; return GINF_OK; }
    int dev = 0;
    GINF_HIP_TRY(hipGetDevice(&dev));
    double *buf = nullptr;
    if (int rc = energy_scratch(dev, &buf)) return rc;
    static_assert(kEnergyBlocks + 1 <= kEnergyScratch, "the energy launch needs kEnergyBlocks partials and a total");
    hipStream_t st = (hipStream_t)stream;
    // k_energy_partial writes every one of the kEnergyBlocks partials and
    // k_energy_final overwrites the total, so nothing has to start at zero.
    hipError_t e = ginf::launch_energy(dtype, data_dev, table_dev, labels_dev, width, height, num_labels, buf,
                                       kEnergyBlocks, buf + kEnergyBlocks, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(energy_host, buf + kEnergyBlocks, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_error(GINF_ERR_HIP, "energy: %s", hipGetErrorString(e));
    if (*energy_host != *energy_host) return set_error(GINF_ERR_INVALID, "a label is outside [0, %d)", num_labels);
    return GINF_OK;
}

// ---------------------------------------------------------------------------
// Batches: many single-level instances advanced together.  The work list is
// grouped by label capacity; each group's half-sweeps are grid-stride launches
// of the kernel built for that capacity, on a graph branch of their own (the
// groups share no state, so they run concurrently and only join at the end).
// The whole decode (relayouts + all half-sweeps) is captured once into a HIP
// graph and replayed.

struct ginf_batch_s {
    int n = 0;
    int dtype = 0, norm = 0, iters = 0;
    std::vector<ginf_bp_plan> plans;
    std::vector<ginf::SweepArgs> args_host[2];
    ginf::SweepArgs *args_dev = nullptr;  // [2][n]
    ginf::BatchItem *items_dev = nullptr;
    int nitems = 0;
    // work-list groups by label capacity and L == capacity or not: items
    // [off, off + count) of items_dev
    static constexpr int kMaxGroups = 12;  // capacities 16, 32, 64, 96, 128, full or not (+ a split largest)
    int ngroups = 0;
    int group_lmax[kMaxGroups] = {}, group_off[kMaxGroups] = {}, group_count[kMaxGroups] = {};
    bool group_full[kMaxGroups] = {};
    std::vector<int> inst_group;  // group of each instance
    hipStream_t group_stream[kMaxGroups] = {};  // capture-only streams of the group branches
    hipEvent_t fork_ev = nullptr, join_ev[kMaxGroups] = {};
    int slots_per_item[4] = {256, 256, 256, 256};  // by label capacity 16, 32, 64, 128
    int blocks_per_cu = 4;
    int num_cu = 0;
    bool seeded = false;  // first half-sweep reads the outboxes stored by ginf_batch_seed
    // ginf_batch_set_predictions: the first half-sweep reads each instance's
    // prediction (args of that half-sweep in pred_args_dev)
    bool predicted = false;
    std::vector<ginf::SweepArgs> pred_args;
    ginf::SweepArgs *pred_args_dev = nullptr;
    bool compact = true;  // 16-bit data costs per instance where they fit (each plan's fit flags)
    bool compact8 = true;  // and 8-bit ones where those fit
    hipStream_t cap = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<const void *> cap_data;
    std::vector<int32_t *> cap_labels;
};

namespace {

int lmax_class(int lmax) { return lmax <= 16 ? 0 : lmax <= 32 ? 1 : lmax <= 64 ? 2 : 3; }

void batch_drop_graph(ginf_batch b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->cap_data.clear();
    b->cap_labels.clear();
}

int batch_build_items(ginf_batch b) {
    std::vector<ginf::BatchItem> items;
    for (int i = 0; i < b->n; ++i) {
        const ginf_bp_plan p = b->plans[i];
        const int ns = p->geom[0].nslots;
        const int per = b->slots_per_item[lmax_class(p->shape.C)];
        for (int s0 = 0; s0 < ns; s0 += per) items.push_back({i, s0, ns - s0 < per ? ns - s0 : per, 0});
    }
    // Grouped by label capacity; within a group the largest instances' items
    // are interleaved with the rest (ordered by item index within the
    // instance) so that every instance is in flight early.
// The largest capacity's instances form two groups (alternate instances):
// that group is the critical path, and two branches give it a larger share
// of the device (64 mixed MRFs: 47.08 ms per replay against 47.44 ms).
#ifndef GINF_BATCH_SPLIT_LARGEST
#define GINF_BATCH_SPLIT_LARGEST 1
#endif
    int cmax = 0;
    for (int i = 0; i < b->n; ++i) cmax = std::max(cmax, b->plans[i]->shape.C);
    auto key = [&](int inst) {  // capacity, then full instances first
        const ginf_bp_plan p = b->plans[inst];
        const int part = (GINF_BATCH_SPLIT_LARGEST && p->shape.C == cmax) ? (inst & 1) : 0;
        return 4 * p->shape.C + 2 * part + (p->desc.num_labels == p->shape.C ? 0 : 1);
    };
    std::stable_sort(items.begin(), items.end(), [&](const ginf::BatchItem &x, const ginf::BatchItem &y) {
        const int kx = key(x.inst), ky = key(y.inst);
        if (kx != ky) return kx < ky;
        const int lx = b->plans[x.inst]->shape.C, ly = b->plans[y.inst]->shape.C;
        const int rx = x.slot0 / b->slots_per_item[lmax_class(lx)];
        const int ry = y.slot0 / b->slots_per_item[lmax_class(ly)];
        return rx < ry;
    });
    b->ngroups = 0;
    int last = -1;
    for (size_t k = 0; k < items.size(); ++k) {
        const int kk = key(items[k].inst);
        if (kk != last) {
            if (b->ngroups == ginf_batch_s::kMaxGroups) return set_error(GINF_ERR_UNSUPPORTED, "too many label capacities");
            last = kk;
            b->group_lmax[b->ngroups] = kk / 4;
            b->group_full[b->ngroups] = (kk & 1) == 0;
            b->group_off[b->ngroups] = (int)k;
            b->group_count[b->ngroups] = 0;
            ++b->ngroups;
        }
        ++b->group_count[b->ngroups - 1];
    }
    b->inst_group.assign(b->n, 0);
    for (int g = 0; g < b->ngroups; ++g)
        for (int k = b->group_off[g]; k < b->group_off[g] + b->group_count[g]; ++k) b->inst_group[items[k].inst] = g;
    if (b->items_dev) { hipFree(b->items_dev); b->items_dev = nullptr; }
    b->nitems = (int)items.size();
    GINF_HIP_TRY(hipMalloc(&b->items_dev, sizeof(ginf::BatchItem) * (items.empty() ? 1 : items.size())));
    if (!items.empty())
        GINF_HIP_TRY(hipMemcpy(b->items_dev, items.data(), sizeof(ginf::BatchItem) * items.size(), hipMemcpyHostToDevice));
    return GINF_OK;
}

int batch_capture(ginf_batch b, const void *const *data_dev, int32_t *const *labels_dev) {
    using namespace ginf;
    batch_drop_graph(b);
    for (int c = 0; c < 2; ++c)
        for (int i = 0; i < b->n; ++i) {
            b->args_host[c][i].labels = labels_dev[i];
        }
    GINF_HIP_TRY(hipMemcpy(b->args_dev, b->args_host[0].data(), sizeof(SweepArgs) * b->n, hipMemcpyHostToDevice));
    GINF_HIP_TRY(hipMemcpy(b->args_dev + b->n, b->args_host[1].data(), sizeof(SweepArgs) * b->n, hipMemcpyHostToDevice));
    if (b->predicted) {
        // half-sweep 0 moves colour 1: its arguments with the prediction as the source
        for (int i = 0; i < b->n; ++i) b->pred_args[i].labels = labels_dev[i];
        GINF_HIP_TRY(hipMemcpy(b->pred_args_dev, b->pred_args.data(), sizeof(SweepArgs) * b->n, hipMemcpyHostToDevice));
    }
    const int N = b->iters;
    hipStream_t st = b->cap;
    for (int k = 0; k < b->ngroups; ++k) {  // branch streams, created on first use
        if (!b->group_stream[k]) GINF_HIP_TRY(hipStreamCreateWithFlags(&b->group_stream[k], hipStreamNonBlocking));
        if (!b->join_ev[k]) GINF_HIP_TRY(hipEventCreateWithFlags(&b->join_ev[k], hipEventDisableTiming));
    }
    GINF_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    auto fail = [&](hipError_t e) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(st, &g);
        if (g) hipGraphDestroy(g);
        return set_error(GINF_ERR_HIP, "batch capture: %s", hipGetErrorString(e));
    };
    hipError_t e = hipEventRecord(b->fork_ev, st);
    if (e != hipSuccess) return fail(e);
    for (int k = 0; k < b->ngroups; ++k)
        if ((e = hipStreamWaitEvent(b->group_stream[k], b->fork_ev, 0)) != hipSuccess) return fail(e);
    // Each instance's data costs (and, for fewer than two half-sweeps, its
    // zero fills and labels) on its group's branch, so that one group's
    // relayouts overlap the other groups' half-sweeps.
    for (int i = 0; i < b->n; ++i) {
        const ginf_bp_plan p = b->plans[i];
        hipStream_t gs = b->group_stream[b->inst_group[i]];
        const StateLayout ly = state_layout(p->shape, p->desc.num_labels);
        auto relayout = [&](int bits, int *flag, const int *only_if) {
            return launch_relayout_pyr(b->dtype, data_dev[i], p->desc.width, p->desc.height, p->desc.num_labels, ly,
                                       p->geom[0], p->lv[0].data[0], p->lv[0].data[1], false, p->geom[0],
                                       p->lv[0].data[0], p->lv[0].data[1], gs, bits <= 16, flag, only_if, bits == 8);
        };
        if (b->compact) {
            // 8-bit costs where the instance's costs fit, else redone in 16 bits, else in T (decided
            // on the device; flag_dev[1] set = not 8-bit, flag_dev[0] set = not 16-bit either)
            if ((e = launch_fill_u32(p->flag_dev, 1, 0u, gs)) != hipSuccess) return fail(e);
            if ((e = launch_fill_u32(p->flag_dev + 1, 1, b->compact8 ? 0u : 1u, gs)) != hipSuccess) return fail(e);
            if (b->compact8 && (e = relayout(8, p->flag_dev + 1, nullptr)) != hipSuccess) return fail(e);
            if ((e = relayout(16, p->flag_dev, b->compact8 ? p->flag_dev + 1 : nullptr)) != hipSuccess) return fail(e);
            if ((e = relayout(32, nullptr, p->flag_dev)) != hipSuccess) return fail(e);
        } else if ((e = relayout(32, nullptr, nullptr)) != hipSuccess) {
            return fail(e);
        }
        if (N <= 1) {
            if (N == 1 && !b->seeded) {
                const int64_t per = (int64_t)ly.C * p->geom[0].nplane * ly.G;
                for (int c = 0; c < 2; ++c)
                    for (int k = 0; k < 4; ++k)
                        if ((e = launch_fill_u32(p->lv[0].msgs[c][k], per, 0u, gs)) != hipSuccess) return fail(e);
            }
            for (int c = 0; c < (N == 1 ? 1 : 2); ++c)
                if ((e = launch_zero_labels(labels_dev[i], p->desc.width, p->desc.height, c, gs)) != hipSuccess)
                    return fail(e);
        }
    }
    const int grid = b->num_cu * b->blocks_per_cu;
    for (int k = 0; k < b->ngroups; ++k) {
        hipStream_t gs = b->group_stream[k];
        for (int it = 0; it < N; ++it) {
            const int color = (it + 1) & 1;
            const bool pred = b->predicted && it == 0;
            const bool zeroin = N >= 2 && it == 0 && !b->seeded && !b->predicted;
            const bool labels = it >= N - 2;
            e = (b->compact ? launch_sweep_batch16 : launch_sweep_batch)(
                b->dtype, b->group_lmax[k], b->group_full[k],
                pred ? b->pred_args_dev : b->args_dev + (size_t)color * b->n, b->items_dev + b->group_off[k],
                b->group_count[k], grid, zeroin, labels, gs, pred);
            if (e != hipSuccess) return fail(e);
        }
        if ((e = hipEventRecord(b->join_ev[k], gs)) != hipSuccess) return fail(e);
        if ((e = hipStreamWaitEvent(st, b->join_ev[k], 0)) != hipSuccess) return fail(e);
    }
    hipGraph_t g = nullptr;
    GINF_HIP_TRY(hipStreamEndCapture(st, &g));
    b->graph = g;
    GINF_HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->cap_data.assign(data_dev, data_dev + b->n);
    b->cap_labels.assign(labels_dev, labels_dev + b->n);
    return GINF_OK;
}

}  // namespace

extern "C" {

int ginf_batch_create(int n, const ginf_bp_desc *descs, const void *const *tables_host, ginf_batch *out) {
    if (!out || !descs || !tables_host || n < 1) return set_error(GINF_ERR_INVALID, "bad batch arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        const ginf_bp_desc &d = descs[i];
        int rc = validate_desc(&d);
        if (rc) return rc;
        if (d.num_levels != 1) return set_error(GINF_ERR_UNSUPPORTED, "batch instance %d: num_levels must be 1", i);
        if (!is_linear(d.sm_model)) return set_error(GINF_ERR_UNSUPPORTED, "batch instance %d: Potts/TL only", i);
        if (d.num_labels > 128) return set_error(GINF_ERR_UNSUPPORTED, "batch instance %d: L > 128", i);
        // the batched kernel is built for label capacities 16, 32, 64, 96 and 128
        if (d.dtype != descs[0].dtype || d.norm != descs[0].norm || d.iters_per_level != descs[0].iters_per_level)
            return set_error(GINF_ERR_INVALID, "batch instances must share dtype, norm and iterations");
    }
    ginf_batch b = new (std::nothrow) ginf_batch_s;
    if (!b) return set_error(GINF_ERR_NOMEM, "host allocation failed");
    b->n = n;
    b->dtype = descs[0].dtype;
    b->norm = descs[0].norm;
    b->iters = descs[0].iters_per_level;
    auto cleanup = [&](int rc) {
        ginf_batch_destroy(b);
        return rc;
    };
    for (int i = 0; i < n; ++i) {
        ginf_bp_desc d = descs[i];
        d.kernel = GINF_KERNEL_QUAD;
        d.lanes_per_node = 0;
        ginf_bp_plan p = nullptr;
        int rc = ginf_bp_plan_create(&d, tables_host[i], &p);
        if (rc) return cleanup(rc);
        b->plans.push_back(p);
        for (int c = 0; c < 2; ++c) {
            ginf::SweepArgs a{};
            const ginf::LevelGeom &g = p->geom[0];
            const ginf::StateLayout ly = ginf::state_layout(p->shape, d.num_labels);
            a.data = p->lv[0].data[c];
            for (int k = 0; k < 4; ++k) { a.out[k] = p->lv[0].msgs[c][k]; a.in[0][k] = p->lv[0].msgs[0][k]; a.in[1][k] = p->lv[0].msgs[1][k]; }
            a.stride = a.in_stride = (int64_t)ginf::kTileSlots * ly.G;
            a.tile = (int64_t)ly.C * ginf::kTileSlots * ly.G;
            a.w = g.w; a.h = g.h; a.whp = g.whp; a.nslots = g.nslots; a.in_whp = g.whp;
            a.color = c;
            a.src = ginf::kSrcSame;
            a.num_labels = d.num_labels;
            a.norm = d.norm;
            a.label_w = d.width;
            a.s_f = p->s_f; a.tau_f = p->tau_f; a.s_i = p->s_i; a.tau_i = p->tau_i;
            a.table = p->table_dev;
            a.arr_bytes = a.data_bytes = a.in_arr_bytes = (uint32_t)p->arr_bytes[0];
            a.plane_bytes = a.in_plane_bytes = (uint32_t)(ginf::kTileSlots * ly.G * 4);
            a.tile_bytes = (uint32_t)(a.tile * 4);
            a.data_plane_bytes = a.plane_bytes / 2;
            a.fit_flag = p->flag_dev;
            a.fit8_flag = p->flag_dev + 1;
            a.out_base = p->lv[0].msgs[c][0];
            a.in_base = p->lv[0].msgs[1 - c][0];
            a.in_span = (uint32_t)(4 * p->arr_bytes[0]);
            a.lmax = p->shape.C;
            b->args_host[c].push_back(a);
        }
    }
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&b->num_cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (b->num_cu <= 0) b->num_cu = 256;
    hipError_t e = hipMalloc(&b->args_dev, sizeof(ginf::SweepArgs) * 2 * n);
    if (e != hipSuccess) return cleanup(set_error(GINF_ERR_NOMEM, "batch args: %s", hipGetErrorString(e)));
    e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(set_error(GINF_ERR_HIP, "batch stream: %s", hipGetErrorString(e)));
    e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming);
    if (e != hipSuccess) return cleanup(set_error(GINF_ERR_HIP, "batch event: %s", hipGetErrorString(e)));
    int rc = batch_build_items(b);
    if (rc) return cleanup(rc);
    *out = b;
    return GINF_OK;
}

int ginf_batch_set_tiles(ginf_batch b, const int *slots_per_item, int blocks_per_cu) {
    if (!b || !slots_per_item) return set_error(GINF_ERR_INVALID, "NULL argument");
    for (int k = 0; k < 4; ++k)
        if (slots_per_item[k] < 64 || slots_per_item[k] % 64) return set_error(GINF_ERR_INVALID, "slots per item must be a positive multiple of 64");
    if (blocks_per_cu < 1 || blocks_per_cu > 64) return set_error(GINF_ERR_INVALID, "blocks_per_cu %d", blocks_per_cu);
    for (int k = 0; k < 4; ++k) b->slots_per_item[k] = slots_per_item[k];
    b->blocks_per_cu = blocks_per_cu;
    batch_drop_graph(b);
    return batch_build_items(b);
}

int ginf_batch_run(ginf_batch b, const void *const *data_dev, int32_t *const *labels_dev, void *stream) {
    if (!b || !data_dev || !labels_dev) return set_error(GINF_ERR_INVALID, "NULL argument");
    for (int i = 0; i < b->n; ++i)
        if (!data_dev[i] || !labels_dev[i]) return set_error(GINF_ERR_INVALID, "NULL pointer for instance %d", i);
    bool same = b->exec && std::equal(b->cap_data.begin(), b->cap_data.end(), data_dev) &&
                std::equal(b->cap_labels.begin(), b->cap_labels.end(), labels_dev);
    if (!same) {
        // The graph is captured on a private stream; order it after prior work on `stream`.
        GINF_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
        int rc = batch_capture(b, data_dev, labels_dev);
        if (rc) return rc;
    }
    GINF_HIP_TRY(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return GINF_OK;
}

int ginf_batch_set_compact_data(ginf_batch b, int enable) {
    if (!b) return set_error(GINF_ERR_INVALID, "NULL batch");
    if (enable < 0 || enable > 2) return set_error(GINF_ERR_INVALID, "compact data mode %d (0, 1 or 2)", enable);
    if (b->compact != (enable != 0) || b->compact8 != (enable == 1)) batch_drop_graph(b);
    b->compact = enable != 0;
    b->compact8 = enable == 1;
    return GINF_OK;
}

int ginf_batch_set_seeded(ginf_batch b, int seeded) {
    if (!b) return set_error(GINF_ERR_INVALID, "NULL batch");
    if (b->seeded != (seeded != 0)) batch_drop_graph(b);
    b->seeded = seeded != 0;
    return GINF_OK;
}

int ginf_batch_set_predictions(ginf_batch b, const float *const *pred_dev, const int *pw, const int *ph, int shift,
                               const float *scale) {
    if (!b) return set_error(GINF_ERR_INVALID, "NULL batch");
    if (!pred_dev) {
        if (b->predicted) batch_drop_graph(b);
        b->predicted = false;
        return GINF_OK;
    }
    if (!pw || !ph || !scale || shift < 0 || shift > 8) return set_error(GINF_ERR_INVALID, "bad prediction arguments");
    if (b->dtype != GINF_F32) return set_error(GINF_ERR_UNSUPPORTED, "predictions need FP32 messages");
    if (b->iters < 2) return set_error(GINF_ERR_UNSUPPORTED, "predictions need two half-sweeps or more");
    std::vector<ginf::SweepArgs> pa((size_t)b->n);
    for (int i = 0; i < b->n; ++i) {
        const int L = b->plans[i]->desc.num_labels;
        if (!pred_dev[i] || pw[i] < 1 || ph[i] < 1 || scale[i] == 0.f)
            return set_error(GINF_ERR_INVALID, "bad prediction for instance %d", i);
        if (L % 4 || ((uintptr_t)pred_dev[i] & 15))
            return set_error(GINF_ERR_UNSUPPORTED, "instance %d: predictions need L %% 4 == 0 and 16-byte alignment", i);
        const size_t bytes = (size_t)pw[i] * ph[i] * 4 * L * sizeof(float);
        if (bytes >= ((size_t)1 << 32)) return set_error(GINF_ERR_UNSUPPORTED, "instance %d: prediction too large", i);
        ginf::SweepArgs a = b->args_host[1][i];  // half-sweep 0 moves colour 1
        a.in_base = pred_dev[i];
        a.in_span = (uint32_t)bytes;
        a.in_plane_bytes = 16;
        a.pred_w = pw[i];
        a.pred_h = ph[i];
        a.pred_shift = shift;
        a.pred_inv_scale = 1.f / scale[i];
        pa[(size_t)i] = a;
    }
    if (!b->pred_args_dev) GINF_HIP_TRY(hipMalloc(&b->pred_args_dev, sizeof(ginf::SweepArgs) * b->n));
    b->pred_args = pa;
    // a captured graph reads the arguments from pred_args_dev at every replay;
    // a new mode needs another kernel instance, so the graph is captured again
    batch_drop_graph(b);
    b->predicted = true;
    return GINF_OK;
}

int ginf_batch_seed(ginf_batch b, int instance, const float *pred_dev, int pw, int ph, int shift, float scale,
                    void *stream) {
    if (!b || instance < 0 || instance >= b->n) return set_error(GINF_ERR_INVALID, "bad instance");
    return ginf_bp_plan_seed(b->plans[instance], 0, pred_dev, pw, ph, shift, scale, stream);
}

int ginf_batch_updates_per_run(ginf_batch b, double *updates) {
    if (!b || !updates) return set_error(GINF_ERR_INVALID, "NULL argument");
    double u = 0.0;
    for (auto p : b->plans) {
        double v = 0.0;
        ginf_bp_plan_updates_per_run(p, &v);
        u += v;
    }
    *updates = u;
    return GINF_OK;
}

int ginf_batch_dump_msgs(ginf_batch b, int instance, void *msgs_host) {
    if (!b || instance < 0 || instance >= b->n) return set_error(GINF_ERR_INVALID, "bad instance");
    return ginf_bp_plan_dump_msgs(b->plans[instance], 0, msgs_host);
}

int ginf_batch_destroy(ginf_batch b) {
    if (!b) return GINF_OK;
    batch_drop_graph(b);
    if (b->cap) hipStreamDestroy(b->cap);
    for (int k = 0; k < ginf_batch_s::kMaxGroups; ++k) {
        if (b->group_stream[k]) hipStreamDestroy(b->group_stream[k]);
        if (b->join_ev[k]) hipEventDestroy(b->join_ev[k]);
    }
    if (b->fork_ev) hipEventDestroy(b->fork_ev);
    if (b->args_dev) hipFree(b->args_dev);
    if (b->pred_args_dev) hipFree(b->pred_args_dev);
    if (b->items_dev) hipFree(b->items_dev);
    for (auto p : b->plans) ginf_bp_plan_destroy(p);
    delete b;
    return GINF_OK;
}

}  // extern "C"
