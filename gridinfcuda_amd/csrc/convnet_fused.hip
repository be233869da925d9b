// The warm-start network's whole inference forward in one launch: six 3x3
// convolutions (L -> 32 -> 32 -> 32 -> 32 -> 32 -> 4L channels, zero padding,
// ReLU and a BF16 store between layers, FP32 output), BF16 MFMA
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
//
// A workgroup owns a tile of TH x TW output pixels of one image.  Its input
// window (the tile plus a 6-pixel ring, one per layer) is staged in LDS once;
// every layer then computes its output over the window shrunk by one more
// pixel per side (the halo the next layers need) from the previous layer's
// LDS buffer into the other one, so no activation leaves the chip and the six
// layers cost one launch instead of six dependent launches, each of which at
// 1/4 resolution is a single memory round trip per wave (about 6 us per
// hidden layer).  The ring pixels are recomputed by the neighbouring tiles.
// Activations of pixels outside the image are stored as 0, which is the next
// layer's zero padding.
//
// Every output element sees the same MFMA sequence as in the layer-by-layer
// kernel (k_conv3x3_pf in conv_net.hip: taps in order, 32-channel steps inside
// each tap, then bias, ReLU and round-to-nearest-even BF16), so the result
// equals ginf_convnet_fwd's layer-by-layer output bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "conv_common.h"
#include "ginf_bp.h"

namespace ginf {
namespace {

constexpr int kLayers = 6;
constexpr int kHid = 32;        // hidden channels
constexpr int kTH = 8;          // output tile rows
constexpr int kTW = 10;         // output tile columns (160x120 -> 16 x 15 = 240 tiles)
constexpr int kWaves = 8;
constexpr int kPitchHid = 80;   // LDS bytes per pixel of a hidden activation (64 + 16: 16 lanes hit distinct banks)

struct NetPtrs {
    const uint16_t *w[kLayers];
    const float *b[kLayers];
};

// LDS bytes per input pixel: CS * 64 data bytes + 16, a multiple of 16 bytes
// (16-byte LDS accesses) whose 4-byte word count leaves the 16 lanes of a
// pixel group (consecutive pixels) on distinct bank quads.
template <int CS>
constexpr int pitch_in() { return CS * 64 + 16; }

// One convolution over an output region of RH x RW pixels whose input region
// (RH + 2) x (RW + 2) is in LDS at `in` (pitch PIN bytes, CS 32-channel
// steps).  The MFMA takes the weights as its first operand and the
// activations as its second, so it yields the transposed output tile: lane
// (m, kg) holds channels 4 kg .. 4 kg + 3 of pixel m, stored with one 8-byte
// (hidden, BF16) or 16-byte (output, FP32) write.  Each product and the order
// of the sums are those of the untransposed product.
// Output co tiles of 16 channels are dealt in chunks of CPC tiles:
// wave w takes chunk w % NCH and the pixel groups w / NCH, w / NCH + NW / NCH,
// ...; a chunk's B fragments stay in registers while its waves walk their
// pixel groups, and each A fragment read from LDS feeds CPC MFMAs.  The
// fragments are loaded (Layer::load) before the barrier that precedes the
// layer, so their latency overlaps the wait for the slowest wave.
//   LAST = false: ReLU, BF16 into `out` (LDS, pitch kPitchHid), 0 outside the image
//   LAST = true:  FP32 into global `gout` (channel count CT * 16), inside pixels only
template <int CS, int CT, int CPC, int RH, int RW, int PIN, bool LAST>
struct Layer {
    static constexpr int CIN = CS * 32;
    static constexpr int NCH = CT / CPC;
    static_assert(CT % CPC == 0 && kWaves % NCH == 0, "co chunks must divide the waves");
    static constexpr int NPIX = RH * RW;
    static constexpr int NG = (NPIX + 15) / 16;
    static constexpr int GSTEP = kWaves / NCH;
    static constexpr int RWI = RW + 2;
    bf16x8 b[CPC][9][CS];
    float bv[CPC][4];  // biases of the lane's channels 4 kg .. 4 kg + 3 of each co tile

    // B fragments of this wave's co chunk straight from the weights in HBM/L2.
    __device__ __forceinline__ void load(const uint16_t *__restrict__ w, const float *__restrict__ bias) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int m = lane & 15, kg = lane >> 4;
        const int chunk = wave % NCH;
#pragma unroll
        for (int j = 0; j < CPC; ++j) {
            const int co = (chunk * CPC + j) * 16 + m;
#pragma unroll
            for (int t = 0; t < 9; ++t)
#pragma unroll
                for (int c = 0; c < CS; ++c) b[j][t][c] = load8(w + ((int64_t)co * 9 + t) * CIN + c * 32 + kg * 8);
#pragma unroll
            for (int i = 0; i < 4; ++i) bv[j][i] = bias[(chunk * CPC + j) * 16 + 4 * kg + i];
        }
    }

    // The same from a copy of the layer's weights in LDS (WStage layout).
    __device__ __forceinline__ void load_lds(const unsigned char *wl, const float *__restrict__ bias) {
        constexpr int R = 9 * CIN / 8 + 1;  // 16-byte chunks per output channel row, one of them padding
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int m = lane & 15, kg = lane >> 4;
        const int chunk = wave % NCH;
#pragma unroll
        for (int j = 0; j < CPC; ++j) {
            const int co = (chunk * CPC + j) * 16 + m;
#pragma unroll
            for (int t = 0; t < 9; ++t)
#pragma unroll
                for (int c = 0; c < CS; ++c)
                    b[j][t][c] = *reinterpret_cast<const bf16x8 *>(wl + (co * R + t * (CIN / 8) + c * 4 + kg) * 16);
#pragma unroll
            for (int i = 0; i < 4; ++i) bv[j][i] = bias[(chunk * CPC + j) * 16 + 4 * kg + i];
        }
    }

    __device__ __forceinline__ void run(const unsigned char *in, unsigned char *out, float *gout, int oy, int ox,
                                        int H, int W, int64_t img_px0) const {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int m = lane & 15, kg = lane >> 4;
        const int chunk = wave % NCH;
        for (int g = wave / NCH; g < NG; g += GSTEP) {
            const int p = g * 16 + m;
            const int pc = p < NPIX ? p : NPIX - 1;  // rows past the region compute a copy and are not stored
            const int r = pc / RW, c0 = pc - (pc / RW) * RW;
            const unsigned char *abase = in + (r * RWI + c0) * PIN + kg * 16;
            // every A fragment of the group is read before the first MFMA, so
            // the LDS latency is paid once per group, not once per tap
            bf16x8 a[9][CS];
#pragma unroll
            for (int t = 0; t < 9; ++t)
#pragma unroll
                for (int c = 0; c < CS; ++c)
                    a[t][c] = *reinterpret_cast<const bf16x8 *>(abase + ((t / 3) * RWI + (t % 3)) * PIN + c * 64);
            f4 acc[CPC];
#pragma unroll
            for (int j = 0; j < CPC; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int t = 0; t < 9; ++t)
#pragma unroll
                for (int c = 0; c < CS; ++c)
#pragma unroll
                    for (int j = 0; j < CPC; ++j)
                        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j][t][c], a[t][c], acc[j], 0, 0, 0);
            // schedule: the group's LDS reads, then its MFMAs (left alone, the
            // scheduler issues two reads ahead and waits before every MFMA pair)
            __builtin_amdgcn_sched_group_barrier(0x100, 9 * CS, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 9 * CS * CPC, 0);
            // lane holds rows 4 * kg + i (channels) of column m (pixel p) of
            // every tile; lanes past the region store nothing
            if (p < NPIX) {
                const int gy = oy + r, gx = ox + c0;
                const bool inside = gy >= 0 && gy < H && gx >= 0 && gx < W;
#pragma unroll
                for (int j = 0; j < CPC; ++j) {
                    const int co0 = (chunk * CPC + j) * 16 + 4 * kg;
                    float v[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[i] = acc[j][i] + bv[j][i];
                    if constexpr (LAST) {
                        if (inside)
                            *reinterpret_cast<f4 *>(gout + (img_px0 + (int64_t)gy * W + gx) * (CT * 16) + co0) =
                                f4{v[0], v[1], v[2], v[3]};
                    } else {
                        uint32_t h[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) h[i] = inside ? f2bf(v[i] > 0.f ? v[i] : 0.f) : 0u;
                        *reinterpret_cast<uint2 *>(out + p * kPitchHid + co0 * 2) =
                            uint2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
                    }
                }
            }
        }
    }
};

// LDS copy of a layer's weights [COUT][9][CIN] (BF16): one row of 16-byte
// chunks per output channel with one chunk of padding, so that the 16 lanes
// reading one chunk column of 16 rows hit distinct banks; rounded up to whole
// 64-chunk wave copies.
template <int COUT, int CIN>
struct WStage {
    static constexpr int R = 9 * CIN / 8 + 1;
    static constexpr int SLOTS = (COUT * R + 63) / 64 * 64;
    static constexpr int BYTES = SLOTS * 16;
};

// Copies a layer's weights into the LDS layout of WStage with
// global_load_lds_dwordx4 (lane l of a copy writes the 16 bytes at
// LDS base + 16 l): the data goes to LDS without passing through registers, and
// the workgroup reads the weights from HBM/L2 once instead of once per wave.
// Padding slots receive chunk 0 again.  Completion is waited for with the
// wave's own vmcnt before the barrier that precedes the reads.
template <int COUT, int CIN>
__device__ __forceinline__ void dma_weights(const uint16_t *__restrict__ w, unsigned char *wl) {
    using S = WStage<COUT, CIN>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = wave; i < S::SLOTS / 64; i += kWaves) {
        const int slot = i * 64 + lane;
        const int row = slot / S::R, k = slot - (slot / S::R) * S::R;
        const int src = (row < COUT && k < S::R - 1) ? row * (S::R - 1) + k : 0;
        __builtin_amdgcn_global_load_lds((const void *)(w + (int64_t)src * 8),
                                         (__attribute__((address_space(3))) void *)(wl + i * 1024), 16, 0, 0);
    }
}

__device__ __forceinline__ void wait_vm() { __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// CS0 = L / 32 input steps, CT5 = 4L / 16 output co tiles.
template <int CS0, int CT5>
__global__ __launch_bounds__(kWaves * 64) void k_convnet6(const uint16_t *__restrict__ x, const NetPtrs net, int H,
                                                          int W, int tiles_x, float *__restrict__ out) {
    constexpr int P0 = pitch_in<CS0>();
    constexpr int RH0 = kTH + 2 * kLayers, RW0 = kTW + 2 * kLayers;
    // Layer 0's weights go through LDS where they fit beside the rest (64
    // labels); with 96 labels each wave loads its co tile's share itself.
    constexpr bool L0_LDS = CS0 <= 2;
    using WH = WStage<kHid, kHid>;
    using W0 = WStage<kHid, CS0 * 32>;
    constexpr int WA_BYTES = L0_LDS && W0::BYTES > WH::BYTES ? W0::BYTES : WH::BYTES;
    __shared__ __attribute__((aligned(16))) unsigned char buf0[RH0 * RW0 * P0];             // input window, layers 1, 3
    __shared__ __attribute__((aligned(16))) unsigned char buf1[(kTH + 10) * (kTW + 10) * kPitchHid];  // layers 0, 2, 4
    __shared__ __attribute__((aligned(16))) unsigned char wA[WA_BYTES];  // weights of layers 0, 2, 4
    __shared__ __attribute__((aligned(16))) unsigned char wB[WH::BYTES];  // weights of layers 1, 3
    const int n = blockIdx.y;
    const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - (blockIdx.x / tiles_x) * tiles_x;
    const int y0 = ty * kTH, x0 = tx * kTW;
    const int64_t img_px0 = (int64_t)n * H * W;
    // Layer k's output region: (kTH + 2 * (5 - k)) x (kTW + 2 * (5 - k)) at
    // (y0 - 5 + k, x0 - 5 + k).
    Layer<CS0, 2, (CS0 <= 2 ? 2 : 1), kTH + 10, kTW + 10, P0, false> l0;
    Layer<1, 2, 2, kTH + 8, kTW + 8, kPitchHid, false> l1;
    Layer<1, 2, 2, kTH + 6, kTW + 6, kPitchHid, false> l2;
    Layer<1, 2, 2, kTH + 4, kTW + 4, kPitchHid, false> l3;
    Layer<1, 2, 2, kTH + 2, kTW + 2, kPitchHid, false> l4;
    Layer<1, CT5, CT5 / kWaves, kTH, kTW, kPitchHid, true> l5;
    if constexpr (L0_LDS) dma_weights<kHid, CS0 * 32>(net.w[0], wA);
    else l0.load(net.w[0], net.b[0]);
    // Stage the input window in 16-byte chunks (zeros outside the image): every
    // load of a thread is issued before its first store.
    {
        constexpr int CH = CS0 * 4;  // 16-byte chunks per pixel
        constexpr int NCHUNK = RH0 * RW0 * CH;
        constexpr int PER = (NCHUNK + kWaves * 64 - 1) / (kWaves * 64);
        const int oy = y0 - kLayers, ox = x0 - kLayers;
        uint4 v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int k = threadIdx.x + u * kWaves * 64;
            const int px = k / CH, ch = k - (k / CH) * CH;
            const int gy = oy + px / RW0, gx = ox + px % RW0;
            v[u] = uint4{0u, 0u, 0u, 0u};
            if (k < NCHUNK && gy >= 0 && gy < H && gx >= 0 && gx < W)
                v[u] = *reinterpret_cast<const uint4 *>(x + (img_px0 + (int64_t)gy * W + gx) * (CS0 * 32) + ch * 8);
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int k = threadIdx.x + u * kWaves * 64;
            const int px = k / CH, ch = k - (k / CH) * CH;
            if (k < NCHUNK) *reinterpret_cast<uint4 *>(buf0 + px * P0 + ch * 16) = v[u];
        }
    }
    wait_vm();
    __syncthreads();
    // Each layer: B fragments from LDS, the next layer's weights copied into
    // the other weight buffer while this layer computes, then the barrier.
    if constexpr (L0_LDS) l0.load_lds(wA, net.b[0]);
    dma_weights<kHid, kHid>(net.w[1], wB);
    l0.run(buf0, buf1, nullptr, y0 - 5, x0 - 5, H, W, img_px0);
    wait_vm();
    __syncthreads();
    l1.load_lds(wB, net.b[1]);
    dma_weights<kHid, kHid>(net.w[2], wA);
    l1.run(buf1, buf0, nullptr, y0 - 4, x0 - 4, H, W, img_px0);
    wait_vm();
    __syncthreads();
    l2.load_lds(wA, net.b[2]);
    dma_weights<kHid, kHid>(net.w[3], wB);
    l2.run(buf0, buf1, nullptr, y0 - 3, x0 - 3, H, W, img_px0);
    wait_vm();
    __syncthreads();
    l3.load_lds(wB, net.b[3]);
    dma_weights<kHid, kHid>(net.w[4], wA);
    l3.run(buf1, buf0, nullptr, y0 - 2, x0 - 2, H, W, img_px0);
    wait_vm();
    __syncthreads();
    l4.load_lds(wA, net.b[4]);
    l5.load(net.w[5], net.b[5]);  // the output layer's co chunks are per wave: loaded from L2, under layer 4
    l4.run(buf0, buf1, nullptr, y0 - 1, x0 - 1, H, W, img_px0);
    __syncthreads();
    l5.run(buf1, nullptr, out, y0, x0, H, W, img_px0);
}

template <int CS0, int CT5>
hipError_t launch_convnet6(const uint16_t *x, const NetPtrs &net, int N, int H, int W, float *out, hipStream_t st) {
    const int tiles_x = (W + kTW - 1) / kTW, tiles_y = (H + kTH - 1) / kTH;
    hipLaunchKernelGGL((k_convnet6<CS0, CT5>), dim3((unsigned)(tiles_x * tiles_y), (unsigned)N), dim3(kWaves * 64),
                       0, st, x, net, H, W, tiles_x, out);
    return hipGetLastError();
}

}  // namespace

// The fused forward for the network shapes it is built for: 6 layers,
// L -> 32 x 5 -> 4L channels with L = 64 or 96.  Returns hipErrorNotSupported
// for any other shape (the caller then runs the layers one by one).
hipError_t launch_convnet_fused(int layers, const int *chans, const uint16_t *const *w, const float *const *b,
                                const uint16_t *x, int N, int H, int W, float *out, hipStream_t st) {
    if (layers != kLayers) return hipErrorNotSupported;
    for (int k = 1; k < kLayers; ++k)
        if (chans[k] != kHid) return hipErrorNotSupported;
    const int L = chans[0];
    if (chans[kLayers] != 4 * L || N < 1 || N > 65535) return hipErrorNotSupported;
    if ((int64_t)((W + kTW - 1) / kTW) * ((H + kTH - 1) / kTH) > 0x7fffffff) return hipErrorNotSupported;
    NetPtrs net{};
    for (int k = 0; k < kLayers; ++k) { net.w[k] = w[k]; net.b[k] = b[k]; }
    if (L == 64) return launch_convnet6<2, 16>(x, net, N, H, W, out, st);
    if (L == 96) return launch_convnet6<3, 24>(x, net, N, H, W, out, st);
    return hipErrorNotSupported;
}

}  // namespace ginf
