// denoise.hip — learned denoiser for path-traced frames (no reference counterpart: the
// reference renders one primary sample per pixel and has no denoiser).
//
// Network: a 5-level convolutional U-Net over 10 per-pixel features (log radiance, albedo,
// normal, log distance; padded to 16 channels).  Encoder level l: two 3x3 convolutions with
// ReLU (widths E[l]) then 2x2 average pooling; level 4 is the bottleneck.  Decoder level l:
// nearest 2x upsampling, concatenation with the encoder's level-l output, two 3x3
// convolutions with ReLU (widths D[l]); a final 3x3 convolution gives linear RGB radiance.
//
// Numerics: activations, weights and activation gradients in BF16 (round to nearest even);
// every convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16 with FP32 accumulation;
// biases, the loss, weight gradients (split-K partial sums reduced in a fixed order, so runs
// are deterministic) and Adam state are FP32.
//
// Kernels:
//   k_conv3x3        forward conv (+bias, ReLU) and backward-data conv: dX = conv(dY, W')
//                    with W'[ci][8-tap][co] = W[co][tap][ci]; its epilogue multiplies by the
//                    ReLU mask of the layer that produced X (X > 0).  Each wave computes 32
//                    pixels x all output channels; A fragments (8 channels of one pixel at
//                    one tap) are 16-byte loads from NHWC tensors, B fragments from the
//                    [co][tap][ci] weight rows.  Weights are the first MFMA operand, so a
//                    lane's outputs are 4 consecutive channels of one pixel.  For inputs of
//                    up to 160 channels the same computation runs from LDS:
//                    k_conv3x3_res (persistent workgroups of 8 waves with the layer's weights
//                    resident in LDS, 16 x 16 pixel tiles with their halo, the next tile
//                    fetched into registers during the current one, a software-pipelined K
//                    loop with no barrier) for up to 96 channels where weights and a tile
//                    fit in 160 KiB, else k_conv3x3_tile (16 x 8 tiles, weights streamed
//                    through LDS in 32-wide K slices).  Both can pool their output (2x2
//                    mean) in the epilogue and read the decoder's upsampled channels from
//                    the lower level in their staging.
//   k_conv3x3_wgrad  dW[co][tap][ci] = sum_p dY[p][co] X[p+tap][ci] over a chunk of pixels:
//                    32-pixel tiles of dY and the shifted X are staged transposed in LDS so
//                    both MFMA operands read 8 consecutive pixels; dbias in the same pass.
//   pooling, upsampling, feature, loss, Adam and weight-layout kernels are element-wise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vt_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kInCh = 16;       // input feature channels (10 used)
constexpr int kOutCh = 3;
constexpr int kLayers = 19;
constexpr int kMaxWidth = 96;   // widest layer (LDS tile of the weight-gradient kernel)
constexpr uint32_t kConvPix = 128;
constexpr uint32_t kPitch = 40;  // LDS row pitch (elements) of 32-pixel transposed tiles
constexpr float kAlbedoEps = 0.01f;  // demodulation: irradiance = radiance / (albedo + eps)
constexpr float kLogMax = 16.0f;     // clamp of the predicted log irradiance

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint32_t b) { return __uint_as_float((b & 0xFFFFu) << 16); }
// two floats to packed BF16 (lo in bits 0-15), round to nearest even: f2bf's result for finite
// values, in one v_cvt_pk_bf16_f32
__device__ __forceinline__ uint32_t pk2bf(float lo, float hi) {
    typedef float f32x2v __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2v{lo, hi}), bf16x2v));
}
__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }
__device__ __forceinline__ uint32_t half16(const u32x4& v, int t) { return (v[t >> 1] >> (16 * (t & 1))) & 0xFFFFu; }

// ---- convolution (forward and backward-data) ------------------------------------------
struct ConvArgs {
    const uint16_t* in;      // N x H x W x in_stride, channels [0, Cin)
    const uint16_t* w;       // [Cout][9 * Cin]
    const float* bias;       // [Cout] or null
    void* out;               // N x H x W x out_stride, channels [out_off, out_off + Cout)
    const uint16_t* mask;    // MASK: N x H x W x mask_stride; out = 0 where mask <= 0
    uint16_t* pool;          // optional: 2x2 mean pool of the BF16 output, N x H/2 x W/2 x Cout
    const uint16_t* up;      // optional: input channels [0, up_C) read from this N x H/2 x W/2 x up_C
                             // tensor at (y/2, x/2) (nearest upsampling), the rest from `in`
    uint32_t up_C;
    uint32_t in_stride, out_stride, out_off, mask_stride;
    uint32_t N, H, W, Cin, Cout;
};

// Epilogue of the LDS-tiled kernels (waves of two 16-pixel rows, weights as the first MFMA
// operand: a lane holds channels 16j + 4g .. 16j + 4g + 3 of pixel r of rows m = 0, 1) for a
// layer whose BF16 output is also pooled (a.pool; H and W even, 4-aligned strides and
// offset).  Rows m = 0, 1 are an even/odd row pair (tile origins are multiples of 8) and lanes
// r, r ^ 1 an even/odd column pair, so the pooled value comes from this lane and its
// neighbour with k_pool_fwd's operations: the four BF16 outputs summed in row-major order,
// times 0.25, rounded to BF16.  Every lane reaches the shuffles; only stores are guarded.
template <int NT, bool RELU>
__device__ __forceinline__ void epilogue_pool(const ConvArgs& a, const f32x4 (&acc)[2][NT], const float (&bias)[NT][4],
                                              uint32_t n, int x, int y0, uint32_t g, uint32_t r) {
    const bool xin = x < (int)a.W, yin = y0 < (int)a.H;
    const size_t img = (size_t)n * a.H * a.W;
    const uint32_t hw = a.W / 2u;
    const size_t q = ((size_t)n * (a.H / 2u) + (size_t)(y0 >> 1)) * hw + (size_t)(x >> 1);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const uint32_t co0 = (uint32_t)j * 16u + 4u * g;
        const bool cin = co0 < a.Cout;
        uint32_t pk[2][2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                v[i] = acc[m][j][i] + bias[j][i];
                if (RELU) v[i] = v[i] > 0.0f ? v[i] : 0.0f;
            }
            pk[m][0] = pk2bf(v[0], v[1]);
            pk[m][1] = pk2bf(v[2], v[3]);
            if (xin && yin && cin) {
                const size_t p = img + (size_t)(y0 + m) * a.W + (size_t)x;
                *reinterpret_cast<uint2*>(reinterpret_cast<uint16_t*>(a.out) + p * a.out_stride + a.out_off + co0) =
                    make_uint2(pk[m][0], pk[m][1]);
            }
        }
        uint32_t nb[2][2];
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int t = 0; t < 2; ++t) nb[m][t] = (uint32_t)__shfl_xor((int)pk[m][t], 1, 64);
        if ((r & 1u) == 0u && xin && yin && cin) {
            float o[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = i >> 1, sh = 16 * (i & 1);
                const float s = ((bf2f(pk[0][t] >> sh) + bf2f(nb[0][t] >> sh)) + bf2f(pk[1][t] >> sh)) +
                                bf2f(nb[1][t] >> sh);
                o[i] = 0.25f * s;
            }
            *reinterpret_cast<uint2*>(a.pool + q * a.Cout + co0) = make_uint2(pk2bf(o[0], o[1]), pk2bf(o[2], o[3]));
        }
    }
}

template <int NT, bool RELU, bool OUTF, bool MASK>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t g = lane >> 4, r = lane & 15u;
    const uint32_t HW = a.H * a.W, P = a.N * HW;
    const uint32_t pbase = blockIdx.x * kConvPix + wave * 32u;
    int py[2], px[2];
    uint32_t img[2];
    bool pv[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const uint32_t p = pbase + m * 16u + r;
        pv[m] = p < P;
        const uint32_t q = pv[m] ? p : 0u;
        const uint32_t n = q / HW, rem = q - n * HW;
        py[m] = (int)(rem / a.W);
        px[m] = (int)(rem - (uint32_t)py[m] * a.W);
        img[m] = n * HW;
    }
    f32x4 acc[2][NT];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[m][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t K = 9u * a.Cin;
    for (uint32_t k0 = 0; k0 < K; k0 += 32u) {
        const uint32_t k = k0 + 8u * g;
        const bool kv = k < K;
        uint32_t tap = 0, ci = 0;
        if (kv) {
            tap = k / a.Cin;
            ci = k - tap * a.Cin;
        }
        const int dy = (int)(tap / 3u) - 1, dx = (int)(tap % 3u) - 1;
        bf16x8 A[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int yy = py[m] + dy, xx = px[m] + dx;
            const bool ok = kv && pv[m] && yy >= 0 && yy < (int)a.H && xx >= 0 && xx < (int)a.W;
            A[m] = ok ? ld8(a.in + (size_t)(img[m] + (uint32_t)yy * a.W + (uint32_t)xx) * a.in_stride + ci) : zero8();
        }
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const uint32_t co = (uint32_t)j * 16u + r;
            const bf16x8 B = (kv && co < a.Cout) ? ld8(a.w + (size_t)co * K + k) : zero8();
            // weights as the first operand: D rows are output channels, columns pixels
            acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B, A[0], acc[0][j], 0, 0, 0);
            acc[1][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B, A[1], acc[1][j], 0, 0, 0);
        }
    }
    // D layout: lane holds output channels 16j + 4g .. 16j + 4g + 3 of pixel r of tile m
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const uint32_t p = pbase + (uint32_t)m * 16u + r;
        if (p >= P) continue;
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t co = (uint32_t)j * 16u + 4u * g + (uint32_t)i;
                if (co >= a.Cout) continue;
                float v = acc[m][j][i] + (a.bias ? a.bias[co] : 0.0f);
                if (RELU) v = v > 0.0f ? v : 0.0f;
                if (MASK && !(bf2f(a.mask[(size_t)p * a.mask_stride + co]) > 0.0f)) v = 0.0f;
                if (OUTF) reinterpret_cast<float*>(a.out)[(size_t)p * a.out_stride + a.out_off + co] = v;
                else reinterpret_cast<uint16_t*>(a.out)[(size_t)p * a.out_stride + a.out_off + co] = f2bf(v);
            }
    }
}

// Tiled form of k_conv3x3 for inputs of at most kTileMaxCin channels: a workgroup (4 waves)
// computes a kTileX x (4 MT) pixel tile of one image.  The input tile with a one-pixel halo is
// staged in LDS once (zeros outside the image), so the nine taps of each pixel read LDS
// instead of fetching the same 16-byte chunks again from L1/L2.  Each wave owns MT rows of 16
// pixels (MT 16-row MFMA tiles) x all output channels, so every weight fragment it reads feeds
// MT MFMAs.  The weights pass through LDS in 32-wide K slices, double-buffered: each slice is
// fetched from global memory into registers while the previous one feeds the MFMAs.  K chunks,
// operands and their order are those of k_conv3x3 (steps past 9 Cin add zero products), so
// every output is the same sum in the same order.
constexpr int kTileX = 16;
constexpr uint32_t kTileMaxCin = 160;   // the sliced kernel (the resident one: kResMaxCin)
constexpr uint32_t kResMaxCin = 96;

// LDS row pitch (elements) of a staged 32-wide K slice of the weights: 80 B rows, so the 16
// rows one MFMA operand reads start in different banks
constexpr uint32_t kBPitch = 40;

template <int NT>
inline size_t conv_tile_lds(int mt, uint32_t cin) {
    return ((size_t)(4 * mt + 2) * (kTileX + 2) * (cin + 8u) + 2u * NT * 16u * kBPitch) * 2u;
}

template <int MT, int NT, bool RELU, bool OUTF, bool MASK>
__global__ __launch_bounds__(256) void k_conv3x3_tile(ConvArgs a) {
    constexpr int TY = 4 * MT;
    extern __shared__ __align__(16) uint16_t tile[];  // (TY + 2) x (kTileX + 2) pixels x cs, then
                                                      // 2 x NT*16 rows x kBPitch weight slices
    const uint32_t cs = a.Cin + 8u;                   // LDS pixel stride (16 B of padding)
    uint16_t* const bsl = tile + (size_t)((TY + 2) * (kTileX + 2)) * cs;
    const uint32_t n = blockIdx.z;
    const int tx0 = (int)blockIdx.x * kTileX, ty0 = (int)blockIdx.y * TY;
    const size_t img = (size_t)n * a.H * a.W;
    // cpp + 1 chunks of 16 B per pixel: its channels, then its padding, set to zeros (K past
    // the end reads the padding, so no select is needed)
    const uint32_t cpp = a.Cin / 8u + 1u;
    const uint32_t total = (uint32_t)((TY + 2) * (kTileX + 2)) * cpp;
    // i / cpp as (i * inv) >> 20: exact while i < 2^20 / cpp (cpp <= 21, i < 6000); i * inv
    // stays below 2^32 since cpp >= 2
    const uint32_t inv = ((1u << 20) + cpp - 1u) / cpp;
    // staging in batches: all loads of a batch are issued before its LDS stores, so a thread
    // waits for one round trip per batch instead of one per 16-byte chunk
    constexpr uint32_t kBatch = 8;
    for (uint32_t i0 = 0; i0 < total; i0 += 256u * kBatch) {
        u32x4 v[kBatch];
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b) {
            const uint32_t i = i0 + threadIdx.x + 256u * b;
            const uint32_t pix = (i * inv) >> 20, c8 = i - pix * cpp;
            const int ly = (int)(pix / (uint32_t)(kTileX + 2)), lx = (int)(pix % (uint32_t)(kTileX + 2));
            const int y = ty0 + ly - 1, x = tx0 + lx - 1;
            v[b] = u32x4{0u, 0u, 0u, 0u};
            if (i < total && c8 + 1u < cpp && y >= 0 && y < (int)a.H && x >= 0 && x < (int)a.W)
                v[b] = *reinterpret_cast<const u32x4*>(
                    (a.up && 8u * c8 < a.up_C)
                        ? a.up + (((size_t)n * (a.H >> 1) + (size_t)(y >> 1)) * (a.W >> 1) + (size_t)(x >> 1)) * a.up_C +
                              8u * c8
                        : a.in + (img + (size_t)y * a.W + (size_t)x) * a.in_stride + 8u * c8);
        }
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b) {
            const uint32_t i = i0 + threadIdx.x + 256u * b;
            const uint32_t pix = (i * inv) >> 20, c8 = i - pix * cpp;
            if (i < total) *reinterpret_cast<u32x4*>(tile + (size_t)pix * cs + 8u * c8) = v[b];
        }
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t g = lane >> 4, r = lane & 15u;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[m][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t K = 9u * a.Cin;
    const uint32_t S = (K + 31u) / 32u;
    // weight slices: slice s holds rows co < NT*16, k = 32s .. 32s+31 (zeros past Cout or K);
    // piece q of a slice is 16 B: row q / 4, k offset 8 (q % 4)
    constexpr uint32_t kPieces = (uint32_t)NT * 64u;
    constexpr uint32_t kPer = (kPieces + 255u) / 256u;
    // loads are unconditional (a clamped address, zeros selected afterwards): conditional
    // loads make the compiler wait for every outstanding load at the join
    // every thread fetches and stores kPer pieces: past kPieces it repeats a piece (the same
    // value to the same place), so nothing is conditional and the compiler keeps the fetch
    // where it is issued instead of sinking it next to the store
    auto piece = [&](uint32_t t) -> uint32_t {
        const uint32_t q = threadIdx.x + 256u * t;
        return q < kPieces ? q : q - kPieces * ((q - kPieces) / kPieces + 1u);
    };
    // (pieces outside the weights are fetched from a clamped address and zeroed at the store,
    // so the wait for the fetch comes at the store)
    auto fetch = [&](uint32_t sl, u32x4 (&dst)[kPer], bool (&ok)[kPer]) {
#pragma unroll
        for (uint32_t t = 0; t < kPer; ++t) {
            const uint32_t q = piece(t);
            const uint32_t co = q >> 2, kk = 32u * sl + 8u * (q & 3u);
            ok[t] = co < a.Cout && kk < K;
            dst[t] = *reinterpret_cast<const u32x4*>(a.w + (ok[t] ? (size_t)co * K + kk : 0u));
        }
    };
    auto store = [&](uint32_t buf, const u32x4 (&src)[kPer], const bool (&ok)[kPer]) {
#pragma unroll
        for (uint32_t t = 0; t < kPer; ++t) {
            const uint32_t q = piece(t);
            *reinterpret_cast<u32x4*>(bsl + (size_t)buf * (NT * 16u * kBPitch) + (q >> 2) * kBPitch + 8u * (q & 3u)) =
                ok[t] ? src[t] : u32x4{0u, 0u, 0u, 0u};
        }
    };
    u32x4 nx[kPer];
    bool nok[kPer];
    fetch(0, nx, nok);
    store(0, nx, nok);
    __syncthreads();
    // k = 32 s + 8 g as (tap, channel): tap = (k * invc) >> 20, exact while k < 2^20 / Cin
    // (k < 9 Cin + 32 <= 1472)
    const uint32_t invc = ((1u << 20) + a.Cin - 1u) / a.Cin;
    const uint32_t rowbase = (uint32_t)MT * wave;
    for (uint32_t sl = 0; sl < S; ++sl) {
        const uint32_t cur = sl & 1u;
        // the next slice (the last one again past the end: a buffer nobody reads)
        fetch(min(sl + 1u, S - 1u), nx, nok);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t k = 32u * sl + 8u * g;
        const uint32_t tap = (k * invc) >> 20;
        const bool kv = tap < 9u;
        const uint32_t tp = kv ? tap : 4u, cc = kv ? k - tap * a.Cin : a.Cin;   // past K: the zero padding
        const uint32_t ty = (tp >= 3u ? 1u : 0u) + (tp >= 6u ? 1u : 0u);
        const int dy = (int)ty - 1, dx = (int)(tp - 3u * ty) - 1;
        bf16x8 A[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int ly = (int)rowbase + m + 1 + dy, lx = (int)r + 1 + dx;
            A[m] = __builtin_bit_cast(
                bf16x8, *reinterpret_cast<const u32x4*>(tile + (size_t)(ly * (kTileX + 2) + lx) * cs + cc));
        }
        const uint16_t* bs = bsl + (size_t)cur * (NT * 16u * kBPitch);
        bf16x8 B[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j)
            B[j] = __builtin_bit_cast(
                bf16x8, *reinterpret_cast<const u32x4*>(bs + ((uint32_t)j * 16u + r) * kBPitch + 8u * g));
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int m = 0; m < MT; ++m)
                acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[j], A[m], acc[m][j], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        store(cur ^ 1u, nx, nok);
        __syncthreads();
    }
    // D layout: lane holds output channels 16j + 4g .. 16j + 4g + 3 of pixel r of its row m.
    // BF16 outputs at 4-channel aligned strides and offsets go out as one 8-byte store per
    // four channels (mask read the same way); otherwise one store per channel
    const int x = tx0 + (int)r;
    if constexpr (MT == 2 && !OUTF && !MASK) {
        if (a.pool) {
            float bias[NT][4];
#pragma unroll
            for (int j = 0; j < NT; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t co = (uint32_t)j * 16u + 4u * g + (uint32_t)i;
                    bias[j][i] = (a.bias && co < a.Cout) ? a.bias[co] : 0.0f;
                }
            epilogue_pool<NT, RELU>(a, acc, bias, n, x, ty0 + (int)rowbase, g, r);
            return;
        }
    }
    if (x >= (int)a.W) return;
    const bool vec = !OUTF && a.Cout % 4u == 0 && a.out_stride % 4u == 0 && a.out_off % 4u == 0 &&
                     (!MASK || a.mask_stride % 4u == 0);
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int y = ty0 + (int)rowbase + m;
        if (y >= (int)a.H) continue;
        const size_t p = img + (size_t)y * a.W + (size_t)x;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const uint32_t co0 = (uint32_t)j * 16u + 4u * g;
            if (co0 >= a.Cout) continue;
            if (vec) {
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    v[i] = acc[m][j][i] + (a.bias ? a.bias[co0 + i] : 0.0f);
                    if (RELU) v[i] = v[i] > 0.0f ? v[i] : 0.0f;
                }
                if (MASK) {
                    const uint2 mk = *reinterpret_cast<const uint2*>(a.mask + p * a.mask_stride + co0);
                    if (!(bf2f(mk.x) > 0.0f)) v[0] = 0.0f;
                    if (!(bf2f(mk.x >> 16) > 0.0f)) v[1] = 0.0f;
                    if (!(bf2f(mk.y) > 0.0f)) v[2] = 0.0f;
                    if (!(bf2f(mk.y >> 16) > 0.0f)) v[3] = 0.0f;
                }
                *reinterpret_cast<uint2*>(reinterpret_cast<uint16_t*>(a.out) + p * a.out_stride + a.out_off + co0) =
                    make_uint2(pk2bf(v[0], v[1]), pk2bf(v[2], v[3]));
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t co = co0 + (uint32_t)i;
                    if (co >= a.Cout) continue;
                    float v = acc[m][j][i] + (a.bias ? a.bias[co] : 0.0f);
                    if (RELU) v = v > 0.0f ? v : 0.0f;
                    if (MASK && !(bf2f(a.mask[p * a.mask_stride + co]) > 0.0f)) v = 0.0f;
                    if (OUTF) reinterpret_cast<float*>(a.out)[p * a.out_stride + a.out_off + co] = v;
                    else reinterpret_cast<uint16_t*>(a.out)[p * a.out_stride + a.out_off + co] = f2bf(v);
                }
            }
        }
    }
}

// Resident-weight form for layers whose weights fit in LDS next to one input tile: persistent
// workgroups of NW waves load the layer's weights into LDS once, as rows of KP = 9 Cin rounded
// up to 64 (zeros past Cout and past 9 Cin) at a pitch of KP + 8 elements (the 16 rows one
// operand reads then start in different banks), and walk the kTileX x (NW MT) pixel tiles of
// the grid.  The next tile's input is fetched into registers while the current tile is
// computed, so the K loop reads only LDS and has no barrier; each K step's tap offset comes
// from a small LDS table.  K chunks, operands and their order are those of k_conv3x3 (steps
// past 9 Cin add zero products), so every output is the same sum in the same order.
template <int MT, int NW>
constexpr uint32_t res_chunks() {   // 16-B chunks of a halo tile per thread, Cin <= 96 (12 + 1 pad)
    return ((uint32_t)(NW * MT + 2) * (kTileX + 2) * 13u + 64u * NW - 1u) / (64u * NW);
}
inline uint32_t res_kp(uint32_t cin) { return (9u * cin + 63u) & ~63u; }
template <int NT>
inline size_t conv_res_lds(int rows, uint32_t cin) {
    return ((size_t)NT * 16u * (res_kp(cin) + 8u) + (size_t)(rows + 2) * (kTileX + 2) * (cin + 8u)) * 2u +
           4u * (res_kp(cin) / 32u + 2u) * 4u;
}

template <int MT, int NW, int NT, bool RELU, bool OUTF, bool MASK>
__global__ __launch_bounds__(64 * NW) void k_conv3x3_res(ConvArgs a, uint32_t ntx, uint32_t nty, uint32_t ntiles) {
    constexpr int TY = NW * MT;
    constexpr uint32_t kThreads = 64u * NW;
    constexpr uint32_t kPix = (uint32_t)(TY + 2) * (kTileX + 2);
    constexpr uint32_t kCh = res_chunks<MT, NW>();
    extern __shared__ __align__(16) uint16_t lds[];
    const uint32_t K = 9u * a.Cin, KP = (K + 63u) & ~63u, WP = KP + 8u, KS = KP / 32u;
    uint16_t* const wl = lds;                                  // NT*16 rows x WP
    uint16_t* const tile = lds + (size_t)NT * 16u * WP;        // kPix pixels x cs
    const uint32_t cs = a.Cin + 8u;
    uint32_t* const offs = reinterpret_cast<uint32_t*>(tile + (size_t)kPix * cs);   // KS x 4 tap offsets
    const uint32_t tid = threadIdx.x;
    // weights -> LDS in batches of 8 pieces of 16 B per thread (row = piece / kq by a
    // multiply-high with ceil(2^32 / kq), exact for these sizes)
    {
        const uint32_t kq = KP / 8u, np = (uint32_t)NT * 16u * kq;
        const uint32_t magic = (uint32_t)(((1ull << 32) + kq - 1u) / kq);
        for (uint32_t i0 = 0; i0 < np; i0 += kThreads * 8u) {
            u32x4 v[8];
#pragma unroll
            for (uint32_t b = 0; b < 8u; ++b) {
                const uint32_t i = i0 + tid + kThreads * b;
                const uint32_t row = __umulhi(i, magic), kk = 8u * (i - row * kq);
                v[b] = u32x4{0u, 0u, 0u, 0u};
                if (i < np && row < a.Cout && kk < K) v[b] = *reinterpret_cast<const u32x4*>(a.w + (size_t)row * K + kk);
            }
#pragma unroll
            for (uint32_t b = 0; b < 8u; ++b) {
                const uint32_t i = i0 + tid + kThreads * b;
                const uint32_t row = __umulhi(i, magic), kk = 8u * (i - row * kq);
                if (i < np) *reinterpret_cast<u32x4*>(wl + (size_t)row * WP + kk) = v[b];
            }
        }
    }
    // tap offsets: entry 4 s + g is the tile offset (elements) of k = 32 s + 8 g relative to the
    // centre tap's pixel; past 9 Cin (and for the two steps after the last, which the pipelined
    // loop reads ahead) it points at the centre pixel's zero padding
    if (tid < 4u * (KS + 2u)) {
        const uint32_t k = 8u * tid;
        const uint32_t tap = k / a.Cin;
        const bool kv = tap < 9u;
        const int dy = kv ? (int)(tap / 3u) - 1 : 0, dx = kv ? (int)(tap % 3u) - 1 : 0;
        offs[tid] = (uint32_t)((dy * (kTileX + 2) + dx) * (int)cs) + (kv ? k - tap * a.Cin : a.Cin);
    }
    // input tiles: cpp = Cin / 8 + 1 chunks per pixel (channels, then the zero padding); chunk
    // i -> pixel (i * inv) >> 20, exact for i < 2^20 / 13
    const uint32_t cpp = a.Cin / 8u + 1u, total = kPix * cpp;
    const uint32_t inv = ((1u << 20) + cpp - 1u) / cpp;
    u32x4 pre[kCh];
    auto origin = [&](uint32_t t, uint32_t& n, int& tx0, int& ty0) {
        const uint32_t tx = t % ntx, rest = t / ntx;
        tx0 = (int)(tx * (uint32_t)kTileX);
        ty0 = (int)((rest % nty) * (uint32_t)TY);
        n = rest / nty;
    };
    auto fetch_tile = [&](uint32_t t) {
        uint32_t n;
        int tx0, ty0;
        origin(t, n, tx0, ty0);
        const size_t img = (size_t)n * a.H * a.W;
#pragma unroll
        for (uint32_t b = 0; b < kCh; ++b) {
            const uint32_t i = tid + kThreads * b;
            const uint32_t pix = (i * inv) >> 20, c8 = i - pix * cpp;
            const int y = ty0 + (int)(pix / (uint32_t)(kTileX + 2)) - 1, x = tx0 + (int)(pix % (uint32_t)(kTileX + 2)) - 1;
            pre[b] = u32x4{0u, 0u, 0u, 0u};
            if (i < total && c8 + 1u < cpp && y >= 0 && y < (int)a.H && x >= 0 && x < (int)a.W)
                pre[b] = *reinterpret_cast<const u32x4*>(
                    (a.up && 8u * c8 < a.up_C)
                        ? a.up + (((size_t)n * (a.H >> 1) + (size_t)(y >> 1)) * (a.W >> 1) + (size_t)(x >> 1)) * a.up_C +
                              8u * c8
                        : a.in + (img + (size_t)y * a.W + (size_t)x) * a.in_stride + 8u * c8);
        }
    };
    uint32_t t = blockIdx.x;
    if (t < ntiles) fetch_tile(t);
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const uint32_t g = lane >> 4, r = lane & 15u;
    const uint32_t rowbase = (uint32_t)MT * wave;
    // this lane's centre pixel of row 0 of its rows, and its weight row
    const uint16_t* const abase = tile + (size_t)((rowbase + 1u) * (kTileX + 2) + r + 1u) * cs;
    const uint32_t rstep = (uint32_t)(kTileX + 2) * cs;
    const uint16_t* const wrow = wl + (size_t)r * WP + 8u * g;
    const bool vec = !OUTF && a.Cout % 4u == 0 && a.out_stride % 4u == 0 && a.out_off % 4u == 0 &&
                     (!MASK || a.mask_stride % 4u == 0);
    // this lane's biases, loaded once: loads in the epilogue would also wait for the next
    // tile's fetch (vector loads complete in order)
    float bias[NT][4];
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t co = (uint32_t)j * 16u + 4u * g + (uint32_t)i;
            bias[j][i] = (a.bias && co < a.Cout) ? a.bias[co] : 0.0f;
        }
    for (; t < ntiles; t += gridDim.x) {
        __syncthreads();   // the previous tile is no longer read (and weights and offsets are written)
#pragma unroll
        for (uint32_t b = 0; b < kCh; ++b) {
            const uint32_t i = tid + kThreads * b;
            const uint32_t pix = (i * inv) >> 20, c8 = i - pix * cpp;
            if (i < total) *reinterpret_cast<u32x4*>(tile + (size_t)pix * cs + 8u * c8) = pre[b];
        }
        __syncthreads();
        uint32_t n;
        int tx0, ty0;
        origin(t, n, tx0, ty0);
        if (t + gridDim.x < ntiles) fetch_tile(t + gridDim.x);
        f32x4 acc[MT][NT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int j = 0; j < NT; ++j) acc[m][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        // software-pipelined by hand: the operands of step s + 1 (and the offset of step s + 2)
        // are read while the MFMAs of step s run; KS is even, reads past the last step hit the
        // zero padding (A) or whatever follows the weight row (B, unused)
        bf16x8 A0[MT], B0[NT], A1[MT], B1[NT];
        auto load = [&](uint32_t ks, uint32_t o, bf16x8 (&A)[MT], bf16x8 (&B)[NT]) {
            const uint16_t* const ap = abase + (int32_t)o;
#pragma unroll
            for (int m = 0; m < MT; ++m)
                A[m] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(ap + (size_t)m * rstep));
#pragma unroll
            for (int j = 0; j < NT; ++j)
                B[j] = __builtin_bit_cast(bf16x8,
                                          *reinterpret_cast<const u32x4*>(wrow + (size_t)j * 16u * WP + 32u * ks));
        };
        auto mma = [&](const bf16x8 (&A)[MT], const bf16x8 (&B)[NT]) {
#pragma unroll
            for (int j = 0; j < NT; ++j)
#pragma unroll
                for (int m = 0; m < MT; ++m)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[j], A[m], acc[m][j], 0, 0, 0);
        };
        load(0u, offs[g], A0, B0);
        uint32_t o1 = offs[4u + g];
        // (scheduling barriers keep each group of reads ahead of the MFMAs it overlaps)
        for (uint32_t ks = 0; ks < KS; ks += 2u) {
            load(ks + 1u, o1, A1, B1);
            const uint32_t o2 = offs[4u * (ks + 2u) + g];
            __builtin_amdgcn_sched_barrier(0);
            mma(A0, B0);
            __builtin_amdgcn_sched_barrier(0);
            load(ks + 2u, o2, A0, B0);
            o1 = offs[4u * (ks + 3u) + g];
            __builtin_amdgcn_sched_barrier(0);
            mma(A1, B1);
            __builtin_amdgcn_sched_barrier(0);
        }
        // D layout: lane holds output channels 16j + 4g .. 16j + 4g + 3 of pixel r of row m
        const int x = tx0 + (int)r;
        if constexpr (MT == 2 && !OUTF && !MASK) {
            if (a.pool) {
                epilogue_pool<NT, RELU>(a, acc, bias, n, x, ty0 + (int)rowbase, g, r);
                continue;
            }
        }
        if (x >= (int)a.W) continue;
        const size_t img = (size_t)n * a.H * a.W;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int y = ty0 + (int)rowbase + m;
            if (y >= (int)a.H) continue;
            const size_t p = img + (size_t)y * a.W + (size_t)x;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const uint32_t co0 = (uint32_t)j * 16u + 4u * g;
                if (co0 >= a.Cout) continue;
                if (vec) {
                    float v[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        v[i] = acc[m][j][i] + bias[j][i];
                        if (RELU) v[i] = v[i] > 0.0f ? v[i] : 0.0f;
                    }
                    if (MASK) {
                        const uint2 mk = *reinterpret_cast<const uint2*>(a.mask + p * a.mask_stride + co0);
                        if (!(bf2f(mk.x) > 0.0f)) v[0] = 0.0f;
                        if (!(bf2f(mk.x >> 16) > 0.0f)) v[1] = 0.0f;
                        if (!(bf2f(mk.y) > 0.0f)) v[2] = 0.0f;
                        if (!(bf2f(mk.y >> 16) > 0.0f)) v[3] = 0.0f;
                    }
                    *reinterpret_cast<uint2*>(reinterpret_cast<uint16_t*>(a.out) + p * a.out_stride + a.out_off + co0) =
                        make_uint2(pk2bf(v[0], v[1]), pk2bf(v[2], v[3]));
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const uint32_t co = co0 + (uint32_t)i;
                        if (co >= a.Cout) continue;
                        float v = acc[m][j][i] + bias[j][i];
                        if (RELU) v = v > 0.0f ? v : 0.0f;
                        if (MASK && !(bf2f(a.mask[p * a.mask_stride + co]) > 0.0f)) v = 0.0f;
                        if (OUTF) reinterpret_cast<float*>(a.out)[p * a.out_stride + a.out_off + co] = v;
                        else reinterpret_cast<uint16_t*>(a.out)[p * a.out_stride + a.out_off + co] = f2bf(v);
                    }
                }
            }
        }
    }
}

// ---- weight gradient ----------------------------------------------------------------------
struct WGradArgs {
    const uint16_t* in;   // forward input, N x H x W x in_stride, channels [0, Cin)
    const uint16_t* g;    // gradient of the pre-activation output, N x H x W x g_stride (g_stride % 8 == 0)
    float* part;          // [chunks][Cout][9 * Cin]
    float* bpart;         // [chunks][Cout]
    uint32_t in_stride, g_stride;
    uint32_t N, H, W, Cin, Cout, chunk_px;
};

// grid (chunks, 9 taps, ceil(Cin / 32)); 256 threads.
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(WGradArgs a) {
    __shared__ __align__(16) uint16_t gT[kMaxWidth][kPitch];
    __shared__ __align__(16) uint16_t xT[32][kPitch];
    const uint32_t chunk = blockIdx.x, tap = blockIdx.y, cib = blockIdx.z;
    const int dy = (int)(tap / 3u) - 1, dx = (int)(tap % 3u) - 1;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t g = lane >> 4, r = lane & 15u;
    const uint32_t HW = a.H * a.W, P = a.N * HW;
    const uint32_t p_begin = chunk * a.chunk_px;
    const uint32_t p_end = min(P, p_begin + a.chunk_px);
    const uint32_t nco16 = (a.Cout + 15u) / 16u;
    const uint32_t c_valid = min(32u, a.Cin - cib * 32u);
    const uint32_t nci16 = (c_valid + 15u) / 16u;
    const uint32_t ntiles = nco16 * nci16;  // <= 12 (Cout <= 96): <= 3 per wave
    const uint32_t cout8 = (a.Cout + 7u) / 8u;
    const bool do_bias = tap == 4u && cib == 0u;
    f32x4 acc[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) acc[s] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    for (uint32_t p0 = p_begin; p0 < p_end; p0 += 32u) {
        for (uint32_t idx = tid; idx < 32u * cout8; idx += 256u) {
            const uint32_t j = idx / cout8, c8 = idx - j * cout8;
            const uint32_t p = p0 + j;
            u32x4 v = {0u, 0u, 0u, 0u};
            if (p < p_end) v = *reinterpret_cast<const u32x4*>(a.g + (size_t)p * a.g_stride + c8 * 8u);
#pragma unroll
            for (int t = 0; t < 8; ++t) gT[c8 * 8u + (uint32_t)t][j] = (uint16_t)half16(v, t);
        }
        if (tid < 128u) {
            const uint32_t j = tid >> 2, o = tid & 3u;
            const uint32_t c = cib * 32u + o * 8u, p = p0 + j;
            u32x4 v = {0u, 0u, 0u, 0u};
            if (p < p_end && c < a.Cin) {
                const uint32_t n = p / HW, rem = p - n * HW;
                const int yy = (int)(rem / a.W) + dy, xx = (int)(rem % a.W) + dx;
                if (yy >= 0 && yy < (int)a.H && xx >= 0 && xx < (int)a.W)
                    v = *reinterpret_cast<const u32x4*>(a.in + (size_t)(n * HW + (uint32_t)yy * a.W + (uint32_t)xx) *
                                                                   a.in_stride + c);
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) xT[o * 8u + (uint32_t)t][j] = (uint16_t)half16(v, t);
        }
        __syncthreads();
        if (do_bias && tid < a.Cout)
            for (int j = 0; j < 32; ++j) bsum += bf2f(gT[tid][j]);
        uint32_t s = 0;
        for (uint32_t ti = wave; ti < ntiles; ti += 4u, ++s) {
            const uint32_t co16 = ti / nci16, c16 = ti - co16 * nci16;
            const bf16x8 A = ld8(&gT[co16 * 16u + r][8u * g]);
            const bf16x8 B = ld8(&xT[c16 * 16u + r][8u * g]);
            if (s == 0) acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[0], 0, 0, 0);
            else if (s == 1) acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[1], 0, 0, 0);
            else acc[2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[2], 0, 0, 0);
        }
        __syncthreads();
    }
    const uint32_t K = 9u * a.Cin;
    uint32_t s = 0;
    for (uint32_t ti = wave; ti < ntiles; ti += 4u, ++s) {
        const uint32_t co16 = ti / nci16, c16 = ti - co16 * nci16;
        const uint32_t c = cib * 32u + c16 * 16u + r;
        const f32x4 v = s == 0 ? acc[0] : (s == 1 ? acc[1] : acc[2]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t co = co16 * 16u + 4u * g + (uint32_t)i;
            if (co < a.Cout && c < a.Cin) a.part[((size_t)chunk * a.Cout + co) * K + tap * a.Cin + c] = v[i];
        }
    }
    if (do_bias && tid < a.Cout) a.bpart[(size_t)chunk * a.Cout + tid] = bsum;
}

// sum of the chunk partials in chunk order (deterministic): W part then bias part
__global__ void k_wgrad_reduce(const float* __restrict__ part, const float* __restrict__ bpart, uint32_t chunks,
                               uint32_t nw, uint32_t nb, float* __restrict__ gw, float* __restrict__ gb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nw) {
        float s = 0.0f;
        for (uint32_t c = 0; c < chunks; ++c) s += part[(size_t)c * nw + i];
        gw[i] = s;
    } else if (i < nw + nb) {
        const uint32_t j = i - nw;
        float s = 0.0f;
        for (uint32_t c = 0; c < chunks; ++c) s += bpart[(size_t)c * nb + j];
        gb[j] = s;
    }
}

// ---- element-wise kernels (8 channels per thread) ----------------------------------------
// out (N x h x w x C) = 2x2 mean of in (N x 2h x 2w, stride, off)
__global__ void k_pool_fwd(const uint16_t* __restrict__ in, uint32_t in_stride, uint32_t in_off,
                           uint16_t* __restrict__ out, uint32_t N, uint32_t h, uint32_t w, uint32_t C) {
    const uint32_t c8n = C / 8u;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * h * w * c8n) return;
    const uint32_t c8 = i % c8n, q = i / c8n;
    const uint32_t x = q % w, y = (q / w) % h, n = q / (w * h);
    const uint32_t W = 2u * w, H = 2u * h;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t d = 0; d < 4; ++d) {
        const uint32_t p = (n * H + 2u * y + (d >> 1)) * W + 2u * x + (d & 1u);
        const u32x4 v = *reinterpret_cast<const u32x4*>(in + (size_t)p * in_stride + in_off + c8 * 8u);
        for (int t = 0; t < 8; ++t) s[t] += bf2f(half16(v, t));
    }
    u32x4 o;
    for (int t = 0; t < 4; ++t)
        o[t] = (uint32_t)f2bf(0.25f * s[2 * t]) | ((uint32_t)f2bf(0.25f * s[2 * t + 1]) << 16);
    *reinterpret_cast<u32x4*>(out + (size_t)q * C + c8 * 8u) = o;
}

// g (N x H x W x C) = (gpool[parent] / 4 + gcat[p, off..]) * (act[p, off..] > 0)
__global__ void k_pool_bwd(const uint16_t* __restrict__ gpool, const uint16_t* __restrict__ gcat,
                           const uint16_t* __restrict__ act, uint32_t stride, uint32_t off, uint16_t* __restrict__ g,
                           uint32_t N, uint32_t H, uint32_t W, uint32_t C) {
    const uint32_t c8n = C / 8u;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * H * W * c8n) return;
    const uint32_t c8 = i % c8n, p = i / c8n;
    const uint32_t x = p % W, y = (p / W) % H, n = p / (W * H);
    const uint32_t parent = (n * (H / 2u) + y / 2u) * (W / 2u) + x / 2u;
    const u32x4 vp = *reinterpret_cast<const u32x4*>(gpool + (size_t)parent * C + c8 * 8u);
    const u32x4 vc = *reinterpret_cast<const u32x4*>(gcat + (size_t)p * stride + off + c8 * 8u);
    const u32x4 va = *reinterpret_cast<const u32x4*>(act + (size_t)p * stride + off + c8 * 8u);
    u32x4 o;
    for (int t = 0; t < 4; ++t) {
        uint32_t h2[2];
        for (int e = 0; e < 2; ++e) {
            const int k = 2 * t + e;
            const float v = 0.25f * bf2f(half16(vp, k)) + bf2f(half16(vc, k));
            h2[e] = bf2f(half16(va, k)) > 0.0f ? f2bf(v) : 0u;
        }
        o[t] = h2[0] | (h2[1] << 16);
    }
    *reinterpret_cast<u32x4*>(g + (size_t)p * C + c8 * 8u) = o;
}

// out (N x 2h x 2w, stride, off) = nearest upsampling of in (N x h x w x C)
__global__ void k_up_fwd(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, uint32_t out_stride,
                         uint32_t out_off, uint32_t N, uint32_t H, uint32_t W, uint32_t C) {
    const uint32_t c8n = C / 8u;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * H * W * c8n) return;
    const uint32_t c8 = i % c8n, p = i / c8n;
    const uint32_t x = p % W, y = (p / W) % H, n = p / (W * H);
    const uint32_t parent = (n * (H / 2u) + y / 2u) * (W / 2u) + x / 2u;
    *reinterpret_cast<u32x4*>(out + (size_t)p * out_stride + out_off + c8 * 8u) =
        *reinterpret_cast<const u32x4*>(in + (size_t)parent * C + c8 * 8u);
}

// g (N x h x w x C) = sum over the 4 children of gcat (N x 2h x 2w, stride, channels [0, C))
__global__ void k_up_bwd(const uint16_t* __restrict__ gcat, uint32_t stride, uint16_t* __restrict__ g, uint32_t N,
                         uint32_t h, uint32_t w, uint32_t C) {
    const uint32_t c8n = C / 8u;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * h * w * c8n) return;
    const uint32_t c8 = i % c8n, q = i / c8n;
    const uint32_t x = q % w, y = (q / w) % h, n = q / (w * h);
    const uint32_t W = 2u * w, H = 2u * h;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t d = 0; d < 4; ++d) {
        const uint32_t p = (n * H + 2u * y + (d >> 1)) * W + 2u * x + (d & 1u);
        const u32x4 v = *reinterpret_cast<const u32x4*>(gcat + (size_t)p * stride + c8 * 8u);
        for (int t = 0; t < 8; ++t) s[t] += bf2f(half16(v, t));
    }
    u32x4 o;
    for (int t = 0; t < 4; ++t) o[t] = (uint32_t)f2bf(s[2 * t]) | ((uint32_t)f2bf(s[2 * t + 1]) << 16);
    *reinterpret_cast<u32x4*>(g + (size_t)q * C + c8 * 8u) = o;
}

// features of an h x w frame into an hp x wp tensor (edge-replicated padding), 16 channels:
// log(1 + max(rad, 0) / (albedo + 0.01)) x3 (demodulated irradiance), albedo x3, normal x3,
// log(1 + distance) / 8, log2(spp) / 8, zeros x5
__global__ void k_features(const float4* __restrict__ rad, const float4* __restrict__ gbuf, uint32_t h, uint32_t w,
                           uint32_t hp, uint32_t wp, float spp_feature, uint16_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hp * wp) return;
    const uint32_t x = min(i % wp, w - 1u), y = min(i / wp, h - 1u);
    const uint32_t s = y * w + x;
    const float4 L = rad[s], A = gbuf[2 * s], Nn = gbuf[2 * s + 1];
    // the albedo the network sees (and remodulates with) is the BF16-rounded one
    const float ar = bf2f(f2bf(A.x)) + kAlbedoEps, ag = bf2f(f2bf(A.y)) + kAlbedoEps, ab = bf2f(f2bf(A.z)) + kAlbedoEps;
    float f[16] = {log1pf(fmaxf(L.x, 0.0f) / ar), log1pf(fmaxf(L.y, 0.0f) / ag), log1pf(fmaxf(L.z, 0.0f) / ab),
                   A.x, A.y, A.z, Nn.x, Nn.y, Nn.z, log1pf(fmaxf(A.w, 0.0f)) * 0.125f, spp_feature,
                   0, 0, 0, 0, 0};
    u32x4 o0, o1;
    for (int t = 0; t < 4; ++t) {
        o0[t] = (uint32_t)f2bf(f[2 * t]) | ((uint32_t)f2bf(f[2 * t + 1]) << 16);
        o1[t] = (uint32_t)f2bf(f[8 + 2 * t]) | ((uint32_t)f2bf(f[9 + 2 * t]) << 16);
    }
    *reinterpret_cast<u32x4*>(out + (size_t)i * 16u) = o0;
    *reinterpret_cast<u32x4*>(out + (size_t)i * 16u + 8u) = o1;
}

__global__ void k_f32_to_bf16(const float* __restrict__ in, uint16_t* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// crop of an h x w frame's RGB target rows (RGBA32F) into N x c x c x 3 at (y0, x0)
struct CropDesc {
    uint32_t frame, y0, x0, pad;
};
struct FrameDesc {
    const uint16_t* feat;   // h x w x 16
    const float* target;    // h x w x 3
    uint32_t h, w;
};
__global__ void k_gather_crops(const FrameDesc* __restrict__ frames, const CropDesc* __restrict__ crops, uint32_t N,
                               uint32_t c, uint16_t* __restrict__ x0, float* __restrict__ tgt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * c * c) return;
    const uint32_t x = i % c, y = (i / c) % c, n = i / (c * c);
    const CropDesc cd = crops[n];
    const FrameDesc fd = frames[cd.frame];
    const size_t s = (size_t)(cd.y0 + y) * fd.w + cd.x0 + x;
    const u32x4* f = reinterpret_cast<const u32x4*>(fd.feat + s * 16u);
    u32x4* o = reinterpret_cast<u32x4*>(x0 + (size_t)i * 16u);
    o[0] = f[0];
    o[1] = f[1];
    for (int k = 0; k < 3; ++k) tgt[(size_t)i * 3u + k] = fd.target[s * 3u + k];
}

__global__ void k_rgba_to_rgb(const float4* __restrict__ in, float* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 v = in[i];
    out[3 * i] = v.x;
    out[3 * i + 1] = v.y;
    out[3 * i + 2] = v.z;
}

// The network predicts a correction y to the input's log irradiance b (feature channels 0..2
// of x): irradiance E = exp(min(b + y, kLogMax)) - 1, radiance a * E with a = albedo + eps
// (feature channels 3..5).  relMSE loss mean((a E - t)^2 / (t^2 + 0.01)) over P pixels x 3
// channels; gy = dLoss/dy (BF16, stride 8, channels 3..7 zero); per-block partial sums.
__device__ __forceinline__ float irradiance(float b, float y, float* dE) {
    const float z = b + y;
    if (z > kLogMax) {
        *dE = 0.0f;
        return expm1f(kLogMax);
    }
    *dE = expf(z);
    return expm1f(z);
}

__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y, const float* __restrict__ t,
                                              const uint16_t* __restrict__ x, uint32_t P, float scale,
                                              uint16_t* __restrict__ gy, float* __restrict__ part) {
    __shared__ float red[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < P) {
        u32x4 o = {0u, 0u, 0u, 0u};
        uint32_t h[3];
        for (int c = 0; c < 3; ++c) {
            const float b = bf2f(x[(size_t)i * 16u + (uint32_t)c]);
            const float a = bf2f(x[(size_t)i * 16u + 3u + (uint32_t)c]) + kAlbedoEps;
            float dE;
            const float E = irradiance(b, y[3 * i + c], &dE);
            const float d = a * E - t[3 * i + c];
            const float den = t[3 * i + c] * t[3 * i + c] + 0.01f;
            l += d * d / den;
            h[c] = f2bf(2.0f * d / den * a * dE * scale);
        }
        o[0] = h[0] | (h[1] << 16);
        o[1] = h[2];
        *reinterpret_cast<u32x4*>(gy + (size_t)i * 8u) = o;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_loss_sum(const float* __restrict__ part, uint32_t n, float scale,
                                                  float* __restrict__ out) {
    __shared__ float red[256];
    float s = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += 256u) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, size_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

struct LayerDev {
    uint32_t cin, cout, cout8;
    uint32_t w_off, b_off;     // in the FP32 parameter vector
    uint32_t wb_off, wt_off;   // BF16 forward weights [co][tap][ci], transposed [ci][8-tap][cout8]
    uint32_t nw;               // cout * 9 * cin
};

// FP32 master -> BF16 forward and transposed (tap-flipped) weights of every layer
__global__ void k_prep_weights(const float* __restrict__ params, const LayerDev* __restrict__ L, uint32_t nl,
                               uint16_t* __restrict__ wb, uint16_t* __restrict__ wt, uint32_t total) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    uint32_t l = 0;
    while (l + 1 < nl && i >= L[l].nw) {
        i -= L[l].nw;
        ++l;
    }
    const LayerDev d = L[l];
    const uint32_t co = i / (9u * d.cin), rem = i - co * 9u * d.cin;
    const uint32_t tap = rem / d.cin, ci = rem - tap * d.cin;
    const uint16_t b = f2bf(params[d.w_off + i]);
    wb[d.wb_off + i] = b;
    wt[d.wt_off + (ci * 9u + (8u - tap)) * d.cout8 + co] = b;
}

// network output (hp x wp x 3 log-irradiance corrections, FP32) -> h x w RGBA32F radiance
// (irradiance remodulated with the albedo of the features x, clamped at 0); alpha < 1 blends
// with the input radiance: alpha * denoised + (1 - alpha) * input
__global__ void k_output(const float* __restrict__ y, const uint16_t* __restrict__ x, uint32_t h, uint32_t w,
                         uint32_t wp, const float4* __restrict__ rad, float alpha, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h * w) return;
    const uint32_t xx = i % w, yy = i / w;
    const size_t q = (size_t)yy * wp + xx;
    const float* v = y + q * 3u;
    const uint16_t* f = x + q * 16u;
    float r[3];
    for (int c = 0; c < 3; ++c) {
        float dE;
        r[c] = fmaxf((bf2f(f[3 + c]) + kAlbedoEps) * irradiance(bf2f(f[c]), v[c], &dE), 0.0f);
    }
    if (alpha < 1.0f) {
        const float4 in = rad[i];
        const float b = 1.0f - alpha;
        r[0] = alpha * r[0] + b * in.x;
        r[1] = alpha * r[1] + b * in.y;
        r[2] = alpha * r[2] + b * in.z;
    }
    out[i] = make_float4(r[0], r[1], r[2], 1.0f);
}

// ---- host side ---------------------------------------------------------------------------
inline uint32_t blocks(size_t n, uint32_t b = 256) { return (uint32_t)((n + b - 1) / b); }
inline uint32_t up8(uint32_t c) { return (c + 7u) & ~7u; }

struct Layer {
    uint32_t cin, cout;
    bool relu;
    LayerDev d;
};

struct Tensor {
    uint16_t* p = nullptr;
    uint32_t C = 0;   // channels (row stride)
};

struct Net {
    bool on = false;
    uint32_t E[5] = {}, D[4] = {}, U[4] = {};
    Layer L[kLayers];
    size_t nparam = 0, nwb = 0, nwt = 0;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    uint16_t *wb = nullptr, *wt = nullptr;
    LayerDev* ldev = nullptr;
    uint64_t adam_t = 0;
    bool weights_dirty = true;
    // activations (capacity in level-0 pixels)
    size_t cap = 0;
    void* act_mem = nullptr;
    Tensor x0, a[4], cat[4], pool[4], b1, b2, d1[4], d2[4];
    float* y = nullptr;
    // gradients
    size_t gcap = 0;
    void* grad_mem = nullptr;
    Tensor gy, gd2[4], gd1[4], gcat[4], gb2, gb1, gpool[4], gskip[4], ga[4];
    float* target = nullptr;
    float* lpart = nullptr;
    float* part = nullptr;
    size_t part_cap = 0;
    float* losses = nullptr;   // device, one per training iteration
    size_t losses_cap = 0;
    // training frames
    std::vector<void*> frame_mem;
    std::vector<FrameDesc> frames;
    FrameDesc* frames_dev = nullptr;
    size_t frames_dev_cap = 0;
    CropDesc* crops_dev = nullptr;
    size_t crops_cap = 0;
    // staging for host <-> device frames
    void* io = nullptr;
    size_t io_cap = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // output blend weights for spp = 2^k, k < nblend (none: the network output alone)
    float blend[32] = {};
    uint32_t nblend = 0;
};

Net net;

// blend weight for spp samples: linear in log2(spp) between the table's entries, clamped to
// its ends; 1 without a table
float blend_alpha(uint32_t spp) {
    if (net.nblend == 0) return 1.0f;
    const float t = log2f((float)(spp ? spp : 1u));
    if (t <= 0.0f) return net.blend[0];
    if (t >= (float)(net.nblend - 1)) return net.blend[net.nblend - 1];
    const uint32_t k = (uint32_t)t;
    const float f = t - (float)k;
    return net.blend[k] + f * (net.blend[k + 1] - net.blend[k]);
}

void set_layer(int i, uint32_t cin, uint32_t cout, bool relu) {
    net.L[i].cin = cin;
    net.L[i].cout = cout;
    net.L[i].relu = relu;
}

void free_act() {
    hipFree(net.act_mem);
    net.act_mem = nullptr;
    net.cap = 0;
}
void free_grad() {
    hipFree(net.grad_mem);
    net.grad_mem = nullptr;
    net.gcap = 0;
}

void free_all() {
    free_act();
    free_grad();
    for (void* p : net.frame_mem) hipFree(p);
    hipFree(net.params); hipFree(net.grads); hipFree(net.adam_m); hipFree(net.adam_v);
    hipFree(net.wb); hipFree(net.wt); hipFree(net.ldev); hipFree(net.part); hipFree(net.losses);
    hipFree(net.frames_dev); hipFree(net.crops_dev); hipFree(net.io);
    if (net.e0) hipEventDestroy(net.e0);
    if (net.e1) hipEventDestroy(net.e1);
    net = Net{};
}

// Bump allocation of tensors out of one device block (256-byte aligned).
struct Bump {
    char* base;
    size_t off = 0;
    bool count_only;
    void* take(size_t bytes) {
        void* p = count_only ? nullptr : base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    }
    Tensor t(size_t px, uint32_t C) {
        Tensor x;
        x.p = static_cast<uint16_t*>(take(px * C * 2));
        x.C = C;
        return x;
    }
};

size_t layout_act(Bump& b, size_t P0) {
    net.x0 = b.t(P0, kInCh);
    for (int l = 0; l < 4; ++l) {
        const size_t P = P0 >> (2 * l);
        net.a[l] = b.t(P, net.E[l]);
        net.cat[l] = b.t(P, net.U[l] + net.E[l]);
        net.pool[l] = b.t(P >> 2, net.E[l]);
        net.d1[l] = b.t(P, net.D[l]);
        net.d2[l] = b.t(P, net.D[l]);
    }
    net.b1 = b.t(P0 >> 8, net.E[4]);
    net.b2 = b.t(P0 >> 8, net.E[4]);
    net.y = static_cast<float*>(b.take(P0 * 3 * 4));
    return b.off;
}

size_t layout_grad(Bump& b, size_t P0) {
    net.gy = b.t(P0, 8);
    for (int l = 0; l < 4; ++l) {
        const size_t P = P0 >> (2 * l);
        net.gd2[l] = b.t(P, net.D[l]);
        net.gd1[l] = b.t(P, net.D[l]);
        net.gcat[l] = b.t(P, net.U[l] + net.E[l]);
        net.gpool[l] = b.t(P >> 2, net.E[l]);
        net.gskip[l] = b.t(P, net.E[l]);
        net.ga[l] = b.t(P, net.E[l]);
    }
    net.gb1 = b.t(P0 >> 8, net.E[4]);
    net.gb2 = b.t(P0 >> 8, net.E[4]);
    net.target = static_cast<float*>(b.take(P0 * 3 * 4));
    net.lpart = static_cast<float*>(b.take((P0 / 256 + 1) * 4));
    return b.off;
}

int ensure_act(size_t P0) {
    if (P0 <= net.cap && net.act_mem) {
        Bump b{static_cast<char*>(net.act_mem), 0, false};
        layout_act(b, P0);
        return 0;
    }
    free_act();
    Bump c{nullptr, 0, true};
    const size_t bytes = layout_act(c, P0);
    if (hipMalloc(&net.act_mem, bytes) != hipSuccess) {
        net.act_mem = nullptr;
        return -1;
    }
    net.cap = P0;
    Bump b{static_cast<char*>(net.act_mem), 0, false};
    layout_act(b, P0);
    return 0;
}

int ensure_grad(size_t P0) {
    if (!(P0 <= net.gcap && net.grad_mem)) {
        free_grad();
        Bump c{nullptr, 0, true};
        const size_t bytes = layout_grad(c, P0);
        if (hipMalloc(&net.grad_mem, bytes) != hipSuccess) {
            net.grad_mem = nullptr;
            return -1;
        }
        net.gcap = P0;
    }
    Bump b{static_cast<char*>(net.grad_mem), 0, false};
    layout_grad(b, P0);
    return 0;
}

// ---- launches ----------------------------------------------------------------------------
// launches one tiled instantiation; workgroups above 64 KiB of LDS are allowed first (once
// per instantiation and size)
template <int MT, int NT, bool RELU, bool OUTF, bool MASK>
int launch_tile(const ConvArgs& a, dim3 tg, hipStream_t s) {
    const size_t lds = conv_tile_lds<NT>(MT, a.Cin);
    static size_t allowed = 65536;
    if (lds > allowed) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3x3_tile<MT, NT, RELU, OUTF, MASK>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return -1;
        allowed = lds;
    }
    hipLaunchKernelGGL((k_conv3x3_tile<MT, NT, RELU, OUTF, MASK>), tg, dim3(256), lds, s, a);
    return 0;
}

// compute units of the current device (queried once)
uint32_t device_cus() {
    static int n = 0;
    if (n == 0 && (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, vt_rt_device()) != hipSuccess || n <= 0))
        n = 256;
    return (uint32_t)n;
}

// resident-weight launch when the weights and a tile fit in LDS and every workgroup gets at
// least two tiles; returns 1 when not applicable
template <int MT, int NW, int NT, bool RELU, bool OUTF, bool MASK>
int launch_res(const ConvArgs& a, hipStream_t s) {
    constexpr uint32_t kMaxLds = 160u * 1024u;
    if (a.Cin > kResMaxCin) return 1;
    const size_t lds = conv_res_lds<NT>(NW * MT, a.Cin);
    if (lds > kMaxLds) return 1;
    const uint32_t ntx = (a.W + kTileX - 1) / kTileX, nty = (a.H + NW * MT - 1u) / (NW * MT);
    const size_t ntiles = (size_t)ntx * nty * a.N;
    const uint32_t per_cu = std::max<uint32_t>(1u, std::min<uint32_t>(8u / NW, (uint32_t)(kMaxLds / lds)));
    const size_t grid = (size_t)device_cus() * per_cu;
    if (ntiles < 2u * grid || ntiles > 0xFFFFFFFFu) return 1;
    static size_t allowed = 65536;
    if (lds > allowed) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3x3_res<MT, NW, NT, RELU, OUTF, MASK>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return -1;
        allowed = lds;
    }
    hipLaunchKernelGGL((k_conv3x3_res<MT, NW, NT, RELU, OUTF, MASK>), dim3((uint32_t)grid), dim3(64 * NW), lds, s, a,
                       ntx, nty, (uint32_t)ntiles);
    return 0;
}

template <bool RELU, bool OUTF, bool MASK>
int launch_conv_nt(uint32_t nt, const ConvArgs& a, hipStream_t s) {
    if (a.Cin <= kTileMaxCin && a.Cin % 8u == 0 && a.in_stride % 8u == 0 && nt <= 6) {
        // resident weights where they fit and the grid has enough tiles, else weight slices
        int rc = 1;
        switch (nt) {
            case 1: rc = launch_res<2, 8, 1, RELU, OUTF, MASK>(a, s); break;
            case 2: rc = launch_res<2, 8, 2, RELU, OUTF, MASK>(a, s); break;
            case 3: rc = launch_res<2, 8, 3, RELU, OUTF, MASK>(a, s); break;
            case 4: rc = launch_res<2, 8, 4, RELU, OUTF, MASK>(a, s); break;
            case 5: rc = launch_res<2, 8, 5, RELU, OUTF, MASK>(a, s); break;
            case 6: rc = launch_res<2, 8, 6, RELU, OUTF, MASK>(a, s); break;
            default: break;
        }
        if (rc <= 0) return rc;
        const dim3 tg((a.W + kTileX - 1) / kTileX, (a.H + 7u) / 8u, a.N);
        switch (nt) {
            case 1: return launch_tile<2, 1, RELU, OUTF, MASK>(a, tg, s);
            case 2: return launch_tile<2, 2, RELU, OUTF, MASK>(a, tg, s);
            case 3: return launch_tile<2, 3, RELU, OUTF, MASK>(a, tg, s);
            case 4: return launch_tile<2, 4, RELU, OUTF, MASK>(a, tg, s);
            case 5: return launch_tile<2, 5, RELU, OUTF, MASK>(a, tg, s);
            case 6: return launch_tile<2, 6, RELU, OUTF, MASK>(a, tg, s);
            default: break;
        }
    }
    const dim3 grid(blocks((size_t)a.N * a.H * a.W, kConvPix)), blk(256);
    switch (nt) {
        case 1: hipLaunchKernelGGL((k_conv3x3<1, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_conv3x3<2, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_conv3x3<3, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_conv3x3<4, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 5: hipLaunchKernelGGL((k_conv3x3<5, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 6: hipLaunchKernelGGL((k_conv3x3<6, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 7: hipLaunchKernelGGL((k_conv3x3<7, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 8: hipLaunchKernelGGL((k_conv3x3<8, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 9: hipLaunchKernelGGL((k_conv3x3<9, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 10: hipLaunchKernelGGL((k_conv3x3<10, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 11: hipLaunchKernelGGL((k_conv3x3<11, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        case 12: hipLaunchKernelGGL((k_conv3x3<12, RELU, OUTF, MASK>), grid, blk, 0, s, a); break;
        default: return -1;
    }
    return 0;
}

// forward conv of layer li: in (Cin channels, stride in.C) -> out (stride out_stride, offset out_off)
// pool (optional, BF16 outputs only): also writes the 2x2 mean pool of the output (N x H/2 x W/2
// x Cout), fused into the epilogue of the LDS-tiled kernels, else by k_pool_fwd afterwards
// up (optional): input channels [0, up.C) are the nearest 2x upsampling of `up` (N x H/2 x W/2),
// read by the LDS-tiled kernels' staging; returns 1 (nothing launched) when the layer does
// not run on those kernels, so the caller materialises the upsampling instead
int conv_fwd(int li, const Tensor& in, void* out, uint32_t out_stride, uint32_t out_off, uint32_t N, uint32_t H,
             uint32_t W, bool outf, hipStream_t s, uint16_t* pool = nullptr, const Tensor* up = nullptr) {
    const Layer& L = net.L[li];
    ConvArgs a{};
    a.in = in.p;
    a.in_stride = in.C;
    a.w = net.wb + L.d.wb_off;
    a.bias = net.params + L.d.b_off;
    a.out = out;
    a.out_stride = out_stride;
    a.out_off = out_off;
    a.N = N; a.H = H; a.W = W;
    a.Cin = L.cin;
    a.Cout = L.cout;
    const uint32_t nt = (L.cout + 15u) / 16u;
    const bool tiled = a.Cin <= kTileMaxCin && a.Cin % 8u == 0 && a.in_stride % 8u == 0 && nt <= 6;
    if (up) {
        if (!tiled || up->C % 8u != 0 || H % 2u != 0 || W % 2u != 0) return 1;
        a.up = up->p;
        a.up_C = up->C;
    }
    if (outf) return launch_conv_nt<false, true, false>(nt, a, s);
    const bool fuse = pool && a.Cin <= kTileMaxCin && a.Cin % 8u == 0 && a.in_stride % 8u == 0 && nt <= 6 &&
                      a.Cout % 4u == 0 && a.out_stride % 4u == 0 && a.out_off % 4u == 0 && H % 2u == 0 && W % 2u == 0;
    a.pool = fuse ? pool : nullptr;
    const int rc = L.relu ? launch_conv_nt<true, false, false>(nt, a, s) : launch_conv_nt<false, false, false>(nt, a, s);
    if (rc == 0 && pool && !fuse) {
        const size_t n8 = (size_t)N * (H / 2) * (W / 2) * (L.cout / 8);
        hipLaunchKernelGGL(k_pool_fwd, dim3(blocks(n8)), dim3(256), 0, s, static_cast<const uint16_t*>(out), out_stride,
                           out_off, pool, N, H / 2, W / 2, L.cout);
    }
    return rc;
}

// backward-data of layer li: g (Cout channels, stride g.C, g.C % 8 == 0) -> dX (Cin channels),
// masked by the forward input X (> 0) when mask is given
int conv_bwd_data(int li, const Tensor& g, const Tensor& dx, const Tensor* mask, uint32_t N, uint32_t H, uint32_t W,
                  hipStream_t s) {
    const Layer& L = net.L[li];
    ConvArgs a{};
    a.in = g.p;
    a.in_stride = g.C;
    a.w = net.wt + L.d.wt_off;
    a.bias = nullptr;
    a.out = dx.p;
    a.out_stride = dx.C;
    a.out_off = 0;
    a.mask = mask ? mask->p : nullptr;
    a.mask_stride = mask ? mask->C : 0;
    a.N = N; a.H = H; a.W = W;
    a.Cin = L.d.cout8;
    a.Cout = L.cin;
    const uint32_t nt = (L.cin + 15u) / 16u;
    return mask ? launch_conv_nt<false, false, true>(nt, a, s) : launch_conv_nt<false, false, false>(nt, a, s);
}

int conv_wgrad(int li, const Tensor& x, const Tensor& g, uint32_t N, uint32_t H, uint32_t W, hipStream_t s) {
    const Layer& L = net.L[li];
    const size_t P = (size_t)N * H * W;
    uint32_t chunk_px = 1024;
    while (chunk_px > 32 && ((P + chunk_px - 1) / chunk_px) * 9u * ((L.cin + 31u) / 32u) < 1024u) chunk_px >>= 1;
    const uint32_t chunks = (uint32_t)((P + chunk_px - 1) / chunk_px);
    const size_t nw = (size_t)L.cout * 9u * L.cin;
    const size_t need = (size_t)chunks * (nw + L.cout);
    if (need > net.part_cap) {
        hipFree(net.part);
        if (hipMalloc(&net.part, need * 4) != hipSuccess) {
            net.part = nullptr;
            net.part_cap = 0;
            return -1;
        }
        net.part_cap = need;
    }
    WGradArgs a{};
    a.in = x.p;
    a.in_stride = x.C;
    a.g = g.p;
    a.g_stride = g.C;
    a.part = net.part;
    a.bpart = net.part + (size_t)chunks * nw;
    a.N = N; a.H = H; a.W = W;
    a.Cin = L.cin;
    a.Cout = L.cout;
    a.chunk_px = chunk_px;
    hipLaunchKernelGGL(k_conv3x3_wgrad, dim3(chunks, 9, (L.cin + 31u) / 32u), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks(nw + L.cout)), dim3(256), 0, s, net.part, a.bpart, chunks,
                       (uint32_t)nw, L.cout, net.grads + L.d.w_off, net.grads + L.d.b_off);
    return 0;
}

void prep_weights(hipStream_t s) {
    if (!net.weights_dirty) return;
    uint32_t total = 0;
    for (auto& L : net.L) total += L.d.nw;
    hipLaunchKernelGGL(k_prep_weights, dim3(blocks(total)), dim3(256), 0, s, net.params, net.ldev, (uint32_t)kLayers,
                       net.wb, net.wt, total);
    net.weights_dirty = false;
}

// forward pass on net.x0 (N x H x W, H and W multiples of 16) -> net.y
// fuse_up: read the decoder's upsampled channels straight from the lower level in the
// convolutions' staging instead of materialising them in cat[l] (inference; training keeps
// cat[l] whole, its weight gradient reads it)
int forward(uint32_t N, uint32_t H, uint32_t W, hipStream_t s, bool fuse_up = false) {
    prep_weights(s);
    int rc = 0;
    Tensor in = net.x0;
    for (int l = 0; l < 4 && rc == 0; ++l) {
        const uint32_t h = H >> l, w = W >> l;
        rc |= conv_fwd(2 * l, in, net.a[l].p, net.a[l].C, 0, N, h, w, false, s);
        rc |= conv_fwd(2 * l + 1, net.a[l], net.cat[l].p, net.cat[l].C, net.U[l], N, h, w, false, s, net.pool[l].p);
        in = net.pool[l];
    }
    const uint32_t h4 = H >> 4, w4 = W >> 4;
    rc |= conv_fwd(8, in, net.b1.p, net.b1.C, 0, N, h4, w4, false, s);
    rc |= conv_fwd(9, net.b1, net.b2.p, net.b2.C, 0, N, h4, w4, false, s);
    Tensor prev = net.b2;
    for (int l = 3; l >= 0 && rc == 0; --l) {
        const uint32_t h = H >> l, w = W >> l;
        const int li = 10 + 2 * (3 - l);
        int fused = 1;
        if (fuse_up) {
            fused = conv_fwd(li, net.cat[l], net.d1[l].p, net.d1[l].C, 0, N, h, w, false, s, nullptr, &prev);
            rc |= fused < 0 ? fused : 0;
        }
        if (fused == 1) {
            const size_t n8 = (size_t)N * h * w * (net.U[l] / 8);
            hipLaunchKernelGGL(k_up_fwd, dim3(blocks(n8)), dim3(256), 0, s, prev.p, net.cat[l].p, net.cat[l].C, 0u, N,
                               h, w, net.U[l]);
            rc |= conv_fwd(li, net.cat[l], net.d1[l].p, net.d1[l].C, 0, N, h, w, false, s);
        }
        rc |= conv_fwd(li + 1, net.d1[l], net.d2[l].p, net.d2[l].C, 0, N, h, w, false, s);
        prev = net.d2[l];
    }
    rc |= conv_fwd(18, net.d2[0], net.y, 3, 0, N, H, W, true, s);
    return rc == 0 && hipGetLastError() == hipSuccess ? 0 : -1;
}

// loss of net.y against net.target and the backward pass -> net.grads; loss into *loss_dev
int backward(uint32_t N, uint32_t H, uint32_t W, float* loss_dev, hipStream_t s) {
    const size_t P0 = (size_t)N * H * W;
    const float scale = 1.0f / (float)(3 * P0);
    const uint32_t lb = blocks(P0);
    hipLaunchKernelGGL(k_loss, dim3(lb), dim3(256), 0, s, net.y, net.target, net.x0.p, (uint32_t)P0, scale,
                       net.gy.p, net.lpart);
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, s, net.lpart, lb, scale, loss_dev);
    int rc = 0;
    // output layer
    rc |= conv_wgrad(18, net.d2[0], net.gy, N, H, W, s);
    rc |= conv_bwd_data(18, net.gy, net.gd2[0], &net.d2[0], N, H, W, s);
    // decoder, level 0 up to 3
    for (int l = 0; l < 4 && rc == 0; ++l) {
        const uint32_t h = H >> l, w = W >> l;
        const int li = 10 + 2 * (3 - l);
        rc |= conv_wgrad(li + 1, net.d1[l], net.gd2[l], N, h, w, s);
        rc |= conv_bwd_data(li + 1, net.gd2[l], net.gd1[l], &net.d1[l], N, h, w, s);
        rc |= conv_wgrad(li, net.cat[l], net.gd1[l], N, h, w, s);
        rc |= conv_bwd_data(li, net.gd1[l], net.gcat[l], &net.cat[l], N, h, w, s);
        // up part of the concatenation -> gradient of the coarser level's output
        const Tensor& gprev = l < 3 ? net.gd2[l + 1] : net.gb2;
        const size_t n8 = (size_t)N * (h / 2) * (w / 2) * (net.U[l] / 8);
        hipLaunchKernelGGL(k_up_bwd, dim3(blocks(n8)), dim3(256), 0, s, net.gcat[l].p, net.gcat[l].C, gprev.p, N,
                           h / 2, w / 2, net.U[l]);
    }
    // bottleneck
    const uint32_t h4 = H >> 4, w4 = W >> 4;
    rc |= conv_wgrad(9, net.b1, net.gb2, N, h4, w4, s);
    rc |= conv_bwd_data(9, net.gb2, net.gb1, &net.b1, N, h4, w4, s);
    rc |= conv_wgrad(8, net.pool[3], net.gb1, N, h4, w4, s);
    rc |= conv_bwd_data(8, net.gb1, net.gpool[3], nullptr, N, h4, w4, s);
    // encoder, level 3 down to 0
    for (int l = 3; l >= 0 && rc == 0; --l) {
        const uint32_t h = H >> l, w = W >> l;
        const size_t n8 = (size_t)N * h * w * (net.E[l] / 8);
        hipLaunchKernelGGL(k_pool_bwd, dim3(blocks(n8)), dim3(256), 0, s, net.gpool[l].p, net.gcat[l].p,
                           net.cat[l].p, net.cat[l].C, net.U[l], net.gskip[l].p, N, h, w, net.E[l]);
        // enc conv 2: input a[l], output (skip) in cat[l]
        Tensor skip_in = net.a[l];
        rc |= conv_wgrad(2 * l + 1, skip_in, net.gskip[l], N, h, w, s);
        rc |= conv_bwd_data(2 * l + 1, net.gskip[l], net.ga[l], &net.a[l], N, h, w, s);
        // enc conv 1: input pool[l-1] (or x0)
        const Tensor& in = l > 0 ? net.pool[l - 1] : net.x0;
        rc |= conv_wgrad(2 * l, in, net.ga[l], N, h, w, s);
        if (l > 0) rc |= conv_bwd_data(2 * l, net.ga[l], net.gpool[l - 1], nullptr, N, h, w, s);
    }
    return rc == 0 && hipGetLastError() == hipSuccess ? 0 : -1;
}

void adam_step(float lr, hipStream_t s) {
    net.adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - powf(b1, (float)net.adam_t), bc2 = 1.0f - powf(b2, (float)net.adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(net.nparam)), dim3(256), 0, s, net.params, net.adam_m, net.adam_v,
                       net.grads, net.nparam, lr, b1, b2, 1e-8f, bc1, bc2);
    net.weights_dirty = true;
}

int ensure_io(size_t bytes) {
    if (bytes <= net.io_cap && net.io) return 0;
    hipFree(net.io);
    net.io = nullptr;
    net.io_cap = 0;
    if (hipMalloc(&net.io, bytes) != hipSuccess) return -1;
    net.io_cap = bytes;
    return 0;
}

uint32_t pcg(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// denoise an h x w frame whose radiance and G-buffer are in device memory; out_dev RGBA32F
float spp_feature(uint32_t spp) { return log2f((float)(spp ? spp : 1u)) * 0.125f; }

int denoise_device(const float4* rad, const float4* gbuf, uint32_t h, uint32_t w, uint32_t spp, float4* out_dev,
                   float* ms, hipStream_t s) {
    const uint32_t hp = (h + 15u) & ~15u, wp = (w + 15u) & ~15u;
    if (ensure_act((size_t)hp * wp) != 0) return -1;
    hipEventRecord(net.e0, s);
    hipLaunchKernelGGL(k_features, dim3(blocks((size_t)hp * wp)), dim3(256), 0, s, rad, gbuf, h, w, hp, wp,
                       spp_feature(spp), net.x0.p);
    if (forward(1, hp, wp, s, true) != 0) return -1;
    hipLaunchKernelGGL(k_output, dim3(blocks((size_t)h * w)), dim3(256), 0, s, net.y, net.x0.p, h, w, wp, rad,
                       blend_alpha(spp), out_dev);
    hipEventRecord(net.e1, s);
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (ms) hipEventElapsedTime(ms, net.e0, net.e1);
    return 0;
}

}  // namespace

extern "C" {

int32_t vt_dn_init(const vt_dn_config* cfg) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!vt_rt_inited() || !cfg) return -1;
    hipSetDevice(vt_rt_device());
    if (net.on) free_all();
    for (int l = 0; l < 5; ++l)
        if (cfg->enc[l] == 0 || cfg->enc[l] % 16 || cfg->enc[l] > (uint32_t)kMaxWidth) {
            vt_rt_set_error("vt_dn_init: widths must be multiples of 16, at most 96");
            return -1;
        }
    for (int l = 0; l < 4; ++l)
        if (cfg->dec[l] == 0 || cfg->dec[l] % 16 || cfg->dec[l] > (uint32_t)kMaxWidth) {
            vt_rt_set_error("vt_dn_init: widths must be multiples of 16, at most 96");
            return -1;
        }
    for (int l = 0; l < 5; ++l) net.E[l] = cfg->enc[l];
    for (int l = 0; l < 4; ++l) net.D[l] = cfg->dec[l];
    net.U[3] = net.E[4];
    for (int l = 2; l >= 0; --l) net.U[l] = net.D[l + 1];
    set_layer(0, kInCh, net.E[0], true);
    set_layer(1, net.E[0], net.E[0], true);
    for (int l = 1; l < 4; ++l) {
        set_layer(2 * l, net.E[l - 1], net.E[l], true);
        set_layer(2 * l + 1, net.E[l], net.E[l], true);
    }
    set_layer(8, net.E[3], net.E[4], true);
    set_layer(9, net.E[4], net.E[4], true);
    for (int l = 3; l >= 0; --l) {
        const int li = 10 + 2 * (3 - l);
        set_layer(li, net.U[l] + net.E[l], net.D[l], true);
        set_layer(li + 1, net.D[l], net.D[l], true);
    }
    set_layer(18, net.D[0], kOutCh, false);
    size_t off = 0, wb = 0, wt = 0;
    LayerDev ld[kLayers];
    for (int i = 0; i < kLayers; ++i) {
        Layer& L = net.L[i];
        L.d.cin = L.cin;
        L.d.cout = L.cout;
        L.d.cout8 = up8(L.cout);
        L.d.nw = L.cout * 9u * L.cin;
        L.d.w_off = (uint32_t)off;
        off += L.d.nw;
        L.d.b_off = (uint32_t)off;
        off += L.cout;
        L.d.wb_off = (uint32_t)wb;
        wb += L.d.nw;
        L.d.wt_off = (uint32_t)wt;
        wt += (size_t)L.cin * 9u * L.d.cout8;
        wb = (wb + 127) & ~(size_t)127;  // keep every layer's rows 16-byte aligned
        wt = (wt + 127) & ~(size_t)127;
        ld[i] = L.d;
    }
    net.nparam = off;
    net.nwb = wb;
    net.nwt = wt;
    hipStream_t s = vt_rt_stream();
    if (hipMalloc(&net.params, off * 4) != hipSuccess || hipMalloc(&net.grads, off * 4) != hipSuccess ||
        hipMalloc(&net.adam_m, off * 4) != hipSuccess || hipMalloc(&net.adam_v, off * 4) != hipSuccess ||
        hipMalloc(&net.wb, wb * 2) != hipSuccess || hipMalloc(&net.wt, wt * 2) != hipSuccess ||
        hipMalloc(&net.ldev, sizeof(ld)) != hipSuccess || hipMalloc(&net.losses, 1024 * 4) != hipSuccess ||
        hipEventCreate(&net.e0) != hipSuccess ||
        hipEventCreate(&net.e1) != hipSuccess) {
        vt_rt_set_error("vt_dn_init: allocation failed");
        free_all();
        return -1;
    }
    hipMemsetAsync(net.params, 0, off * 4, s);
    hipMemsetAsync(net.grads, 0, off * 4, s);
    hipMemsetAsync(net.adam_m, 0, off * 4, s);
    hipMemsetAsync(net.adam_v, 0, off * 4, s);
    hipMemsetAsync(net.wb, 0, wb * 2, s);
    hipMemsetAsync(net.wt, 0, wt * 2, s);  // padding columns (co >= Cout) stay zero
    hipMemcpyAsync(net.ldev, ld, sizeof(ld), hipMemcpyHostToDevice, s);
    if (hipStreamSynchronize(s) != hipSuccess) {
        free_all();
        return -1;
    }
    net.losses_cap = 1024;
    net.on = true;
    net.weights_dirty = true;
    return 0;
}

int32_t vt_dn_layers(uint32_t* out4, uint32_t max_layers) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on) return -1;
    for (uint32_t i = 0; i < (uint32_t)kLayers && i < max_layers && out4; ++i) {
        out4[4 * i] = net.L[i].cin;
        out4[4 * i + 1] = net.L[i].cout;
        out4[4 * i + 2] = net.L[i].d.w_off;
        out4[4 * i + 3] = net.L[i].d.b_off;
    }
    return kLayers;
}

uint64_t vt_dn_param_count(void) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    return net.on ? net.nparam : 0;
}

int32_t vt_dn_set_params(const float* params, uint64_t n) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on || !params || n != net.nparam) return -1;
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    hipMemcpyAsync(net.params, params, n * 4, hipMemcpyHostToDevice, s);
    hipMemsetAsync(net.adam_m, 0, n * 4, s);
    hipMemsetAsync(net.adam_v, 0, n * 4, s);
    net.adam_t = 0;
    net.weights_dirty = true;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

int32_t vt_dn_get_params(float* params, uint64_t n) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on || !params || n != net.nparam) return -1;
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    hipMemcpyAsync(params, net.params, n * 4, hipMemcpyDeviceToHost, s);
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

int32_t vt_dn_grad(const float* x, const float* target, uint32_t n, uint32_t h, uint32_t w, float* y, float* grad,
                   float* loss) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on || !x || !target || n == 0 || h == 0 || w == 0 || h % 16 || w % 16) {
        vt_rt_set_error("vt_dn_grad: bad arguments (h and w must be multiples of 16)");
        return -1;
    }
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    const size_t P0 = (size_t)n * h * w;
    if (ensure_act(P0) != 0 || ensure_grad(P0) != 0 || ensure_io(P0 * kInCh * 4) != 0) {
        vt_rt_set_error("vt_dn_grad: allocation failed");
        return -1;
    }
    hipMemcpyAsync(net.io, x, P0 * kInCh * 4, hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(P0 * kInCh)), dim3(256), 0, s, static_cast<const float*>(net.io),
                       net.x0.p, P0 * kInCh);
    hipMemcpyAsync(net.target, target, P0 * 3 * 4, hipMemcpyHostToDevice, s);
    if (forward(n, h, w, s) != 0 || backward(n, h, w, net.losses, s) != 0) {
        vt_rt_set_error("vt_dn_grad: launch failed");
        return -1;
    }
    if (y) hipMemcpyAsync(y, net.y, P0 * 3 * 4, hipMemcpyDeviceToHost, s);
    if (grad) hipMemcpyAsync(grad, net.grads, net.nparam * 4, hipMemcpyDeviceToHost, s);
    if (loss) hipMemcpyAsync(loss, net.losses, 4, hipMemcpyDeviceToHost, s);
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

int32_t vt_dn_features(const float* rad_rgba, const float* gbuf8, uint32_t h, uint32_t w, uint32_t spp, float* out16) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on || !rad_rgba || !gbuf8 || !out16 || h == 0 || w == 0) return -1;
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    const size_t npx = (size_t)h * w;
    if (ensure_io(npx * (16 + 32 + 32 + 64)) != 0) return -1;
    char* io = static_cast<char*>(net.io);
    float4* rad = reinterpret_cast<float4*>(io);
    float4* gb = reinterpret_cast<float4*>(io + npx * 16);
    uint16_t* f = reinterpret_cast<uint16_t*>(io + npx * 48);
    hipMemcpyAsync(rad, rad_rgba, npx * 16, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(gb, gbuf8, npx * 32, hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(k_features, dim3(blocks(npx)), dim3(256), 0, s, rad, gb, h, w, h, w, spp_feature(spp), f);
    std::vector<uint16_t> hb(npx * 16);
    hipMemcpyAsync(hb.data(), f, npx * 32, hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    for (size_t i = 0; i < hb.size(); ++i) {
        const uint32_t u = (uint32_t)hb[i] << 16;
        std::memcpy(&out16[i], &u, 4);
    }
    return 0;
}

int32_t vt_dn_add_frame(const float* rad_rgba, const float* gbuf8, const float* target_rgba, uint32_t h, uint32_t w,
                        uint32_t spp) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on || !rad_rgba || !gbuf8 || !target_rgba || h < 16 || w < 16) return -1;
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    const size_t npx = (size_t)h * w;
    if (ensure_io(npx * 64) != 0) return -1;
    void* mem = nullptr;
    const size_t fbytes = (npx * 32 + 255) & ~(size_t)255;
    if (hipMalloc(&mem, fbytes + npx * 12) != hipSuccess) {
        vt_rt_set_error("vt_dn_add_frame: allocation failed");
        return -1;
    }
    char* io = static_cast<char*>(net.io);
    float4* rad = reinterpret_cast<float4*>(io);
    float4* gb = reinterpret_cast<float4*>(io + npx * 16);
    FrameDesc fd;
    fd.feat = static_cast<const uint16_t*>(mem);
    fd.target = reinterpret_cast<const float*>(static_cast<char*>(mem) + fbytes);
    fd.h = h;
    fd.w = w;
    hipMemcpyAsync(rad, rad_rgba, npx * 16, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(gb, gbuf8, npx * 32, hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(k_features, dim3(blocks(npx)), dim3(256), 0, s, rad, gb, h, w, h, w, spp_feature(spp),
                       static_cast<uint16_t*>(mem));
    hipMemcpyAsync(rad, target_rgba, npx * 16, hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(k_rgba_to_rgb, dim3(blocks(npx)), dim3(256), 0, s, rad, const_cast<float*>(fd.target),
                       (uint32_t)npx);
    if (hipStreamSynchronize(s) != hipSuccess) {
        hipFree(mem);
        return -1;
    }
    net.frame_mem.push_back(mem);
    net.frames.push_back(fd);
    return (int32_t)net.frames.size() - 1;
}

int32_t vt_dn_clear_frames(void) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on) return -1;
    hipSetDevice(vt_rt_device());
    hipStreamSynchronize(vt_rt_stream());
    for (void* p : net.frame_mem) hipFree(p);
    net.frame_mem.clear();
    net.frames.clear();
    return 0;
}

int32_t vt_dn_train(uint32_t iters, uint32_t batch, uint32_t crop, float lr, uint32_t seed, float* losses) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on || net.frames.empty() || iters == 0 || batch == 0 || crop < 16 || crop % 16) {
        vt_rt_set_error("vt_dn_train: no frames or bad arguments (crop must be a multiple of 16)");
        return -1;
    }
    for (auto& f : net.frames)
        if (f.h < crop || f.w < crop) {
            vt_rt_set_error("vt_dn_train: a frame is smaller than the crop");
            return -1;
        }
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    const size_t P0 = (size_t)batch * crop * crop;
    if (ensure_act(P0) != 0 || ensure_grad(P0) != 0) {
        vt_rt_set_error("vt_dn_train: allocation failed");
        return -1;
    }
    if (net.losses_cap < iters) {
        hipFree(net.losses);
        net.losses = nullptr;
        net.losses_cap = 0;
        if (hipMalloc(&net.losses, (size_t)iters * 4) != hipSuccess) return -1;
        net.losses_cap = iters;
    }
    if (net.frames_dev_cap < net.frames.size()) {
        hipFree(net.frames_dev);
        net.frames_dev = nullptr;
        if (hipMalloc(&net.frames_dev, net.frames.size() * sizeof(FrameDesc)) != hipSuccess) return -1;
        net.frames_dev_cap = net.frames.size();
    }
    hipMemcpyAsync(net.frames_dev, net.frames.data(), net.frames.size() * sizeof(FrameDesc), hipMemcpyHostToDevice, s);
    const size_t ncrop = (size_t)iters * batch;
    if (net.crops_cap < ncrop) {
        hipFree(net.crops_dev);
        net.crops_dev = nullptr;
        if (hipMalloc(&net.crops_dev, ncrop * sizeof(CropDesc)) != hipSuccess) return -1;
        net.crops_cap = ncrop;
    }
    // crop positions: a seeded PCG stream, uniform over frames and positions
    std::vector<CropDesc> crops(ncrop);
    uint32_t st = pcg(seed ^ 0x9E3779B9u);
    for (auto& c : crops) {
        st = pcg(st);
        c.frame = st % (uint32_t)net.frames.size();
        const FrameDesc& f = net.frames[c.frame];
        st = pcg(st);
        c.y0 = st % (f.h - crop + 1u);
        st = pcg(st);
        c.x0 = st % (f.w - crop + 1u);
        c.pad = 0;
    }
    hipMemcpyAsync(net.crops_dev, crops.data(), ncrop * sizeof(CropDesc), hipMemcpyHostToDevice, s);
    if (hipStreamSynchronize(s) != hipSuccess) return -1;  // the host vector must outlive the copy
    const uint32_t warm = std::min<uint32_t>(100u, iters / 10u + 1u);
    for (uint32_t it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_gather_crops, dim3(blocks(P0)), dim3(256), 0, s, net.frames_dev,
                           net.crops_dev + (size_t)it * batch, batch, crop, net.x0.p, net.target);
        if (forward(batch, crop, crop, s) != 0 || backward(batch, crop, crop, net.losses + it, s) != 0) {
            vt_rt_set_error("vt_dn_train: launch failed");
            return -1;
        }
        // linear warm-up, then cosine decay to 5% of lr
        const float t = (float)it / (float)iters;
        float lr_t = lr * (0.05f + 0.95f * 0.5f * (1.0f + cosf(3.14159265f * t)));
        if (it < warm) lr_t *= (float)(it + 1) / (float)warm;
        adam_step(lr_t, s);
    }
    if (losses) hipMemcpyAsync(losses, net.losses, (size_t)iters * 4, hipMemcpyDeviceToHost, s);
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

int32_t vt_dn_denoise(const float* rad_rgba, const float* gbuf8, uint32_t h, uint32_t w, uint32_t spp,
                      float* out_rgba, float* gpu_ms) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on || !rad_rgba || !gbuf8 || !out_rgba || h == 0 || w == 0) return -1;
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    const size_t npx = (size_t)h * w;
    if (ensure_io(npx * 64) != 0) return -1;
    char* io = static_cast<char*>(net.io);
    float4* rad = reinterpret_cast<float4*>(io);
    float4* gb = reinterpret_cast<float4*>(io + npx * 16);
    float4* out = reinterpret_cast<float4*>(io + npx * 48);
    hipMemcpyAsync(rad, rad_rgba, npx * 16, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(gb, gbuf8, npx * 32, hipMemcpyHostToDevice, s);
    if (denoise_device(rad, gb, h, w, spp, out, gpu_ms, s) != 0) {
        vt_rt_set_error("vt_dn_denoise: failed");
        return -1;
    }
    hipMemcpyAsync(out_rgba, out, npx * 16, hipMemcpyDeviceToHost, s);
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

int32_t vt_dn_denoise_frame(float* out_rgba, float* gpu_ms) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    const float4 *rad = nullptr, *gb = nullptr;
    uint32_t w = 0, h = 0, spp = 1;
    if (!net.on || !out_rgba || !vt_rt_frame(&rad, &gb, &w, &h, &spp)) {
        vt_rt_set_error("vt_dn_denoise_frame: no network or no frame with a G-buffer");
        return -1;
    }
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    const size_t npx = (size_t)h * w;
    if (ensure_io(npx * 16) != 0) return -1;
    float4* out = static_cast<float4*>(net.io);
    if (denoise_device(rad, gb, h, w, spp, out, gpu_ms, s) != 0) {
        vt_rt_set_error("vt_dn_denoise_frame: failed");
        return -1;
    }
    hipMemcpyAsync(out_rgba, out, npx * 16, hipMemcpyDeviceToHost, s);
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

int32_t vt_dn_set_blend(const float* alpha, uint32_t n) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!net.on || n > 32 || (n && !alpha)) return -1;
    for (uint32_t k = 0; k < n; ++k)
        if (!(alpha[k] >= 0.0f && alpha[k] <= 1.0f)) return -1;
    for (uint32_t k = 0; k < n; ++k) net.blend[k] = alpha[k];
    net.nblend = n;
    return 0;
}

void vt_dn_shutdown(void) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    vt_dn_release();
}

}  // extern "C"

// Frees the denoiser; the caller holds the runtime lock (cleanup() calls it too).
void vt_dn_release() {
    if (!net.on && !net.params) return;
    hipSetDevice(vt_rt_device());
    hipStreamSynchronize(vt_rt_stream());
    free_all();
}
