// dn_kernels.h -- gfx950 kernels of the AO / bent-normal denoiser (a 2-D
// U-Net over the baked map): 3x3 convolutions as implicit GEMMs on the bf16
// matrix cores (v_mfma_f32_16x16x32_bf16, fp32 accumulation), their weight
// gradients, 2x2 average pooling, the 1x1 output head with the L1 loss, and
// Adam.
//
// Tensors are NHWC: pixel-major, channels contiguous, bf16 stored as uint16_t
// bit patterns.  A 16-pixel x 32-channel MFMA operand is then one 16-byte
// read per lane: lane l takes pixel (l & 15) and channels 8*(l >> 4) .. +7.
// Every convolution input has a multiple of 32 channels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fornos {
namespace dn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;  // 4 waves

__device__ __host__ inline float bf2f(uint16_t u)
{
    return __builtin_bit_cast(float, (uint32_t)u << 16);
}

// round to nearest even; NaN stays NaN
__device__ __host__ inline uint16_t f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline bool bf_pos(uint16_t u) { return !(u & 0x8000u) && u != 0; }

__device__ inline bf16x8 load8(const uint16_t* p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ inline bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

// One 3x3, stride-1, zero-padded convolution.  Input channels [0, c1) come
// from `a` (at half resolution, nearest-upsampled, when up_a), channels
// [c1, c1 + c2) from `b`: the decoder's concat is never materialised.
// Output channels [0, split) go to out (split channels per pixel), channels
// [split, cout) to out2; a mask tensor zeroes an output where the mask value
// is not positive (the ReLU derivative in the backward pass); acc2 adds into
// out2 instead of overwriting it.  The same kernel computes the backward
// pass with respect to the input, with weights flipped and transposed.
//
// Forward-pass fusions (bf16 outputs, split == cout, no mask):
// - in_f32: the input is the network's fp32 image with in_c <= 16 channels
//   (a multiple of 4), converted to bf16 while staged as k_pack_input would,
//   and the weights come packed by k_pack_first ([cout][9 x 16 + 16], `a` is
//   not read): the first layer, whose other input channels are zero;
// - pool: the 2x2 average of the bf16 outputs is written there as well, as
//   k_pool computes it from them (h, w even);
// - head_out: the 1x1 head of k_head on the bf16 outputs (cout = 32, one
//   block holds a pixel's channels), residual on head_in's channels 0..3;
//   out may then be null (the activations are not stored).
struct ConvArgs {
    const uint16_t* a;
    const uint16_t* b;
    const uint16_t* wts;    // [cout][9][c1 + c2]
    const float* bias;      // [cout] or null
    void* out;              // bf16 or fp32 [pixels][split], or null (head_out only)
    void* out2;             // bf16 [pixels][cout - split]
    const uint16_t* mask1;  // [pixels][split] or null
    const uint16_t* mask2;  // [pixels][cout - split] or null
    const float* in_f32;    // [pixels][in_c] or null
    uint16_t* pool;         // [pixels / 4][cout] or null
    const float* head_in;   // [pixels][head_cin]
    const float* head_w;    // [4][32]
    const float* head_b;    // [4]
    float* head_out;        // [pixels][4] or null
    uint32_t c1, c2, cout, split, n, h, w;
    uint32_t in_c, head_cin;
    int up_a, relu, out_f32, acc2;
};

// LDS-tiled implicit GEMM.  A block computes a TH x 32 tile of output
// pixels (TH = 2 * MT rows, MT / 2 rows per wave) for NB = 16 * NT output
// channels.  For each 32-channel slice of the input it stages the tile plus
// its one-pixel halo ((TH + 2) x 34 pixels, 64 bytes each) and the slice's
// weights for the 9 taps ([tap][cout][32 channels]) in LDS once; the 9 taps
// then read their shifted 16-pixel operands from LDS.  A slice's global loads
// are all issued before any is stored to LDS (one latency per slice, not one
// per load), and the next slice's loads are in flight while the current
// slice's MFMAs run.  Pixels outside the image read as zero (the zero
// padding) and are not written.
//
// LDS layout: 16-byte word q (channels 8q .. 8q + 7) of row p (a pixel, or a
// tap's output channel) sits at word 4p + q.
__device__ inline int lds_word(int p, int q) { return 4 * p + q; }

// FUSE: which of the fusions the kernel is built with (kFuse*; the
// arguments select them too), so the kernels without them keep their
// register allocation.
enum { kFuseF32In = 1, kFusePool = 2, kFuseHead = 4, kOneSlice = 8 };  // kOneSlice: cin = 32
// Waves per SIMD the kernel is held to: 5 for the first layer (21 KB of
// LDS per block; 6 spills), 4 for the other single-slice 32-channel layers
// (no next-slice prefetch registers, and their LDS allows 4 blocks per CU),
// else the compiler's choice (64 output channels: LDS allows 2 blocks per CU
// either way).  Forward 2.006 -> 1.91 ms.
template <int NT, int FUSE>
constexpr int conv_waves()
{
    return (FUSE & kFuseF32In) ? 5 : (NT == 2 && (FUSE & kOneSlice)) ? 4 : 1;
}
template <int MT, int NT, int FUSE = 0>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(conv_waves<NT, FUSE>()))) void k_conv3x3(
    ConvArgs A, uint32_t tiles_x, uint32_t tiles_y)
{
    constexpr int TW = 32, TH = 2 * MT, IW = TW + 2, IH = TH + 2, NB = 16 * NT;
    // [pixel][4 x 8 channels] and [tap][cout][4 x 8 channels]; after the last
    // slice, the block's bf16 outputs [pixel][NB channels] (forward pass)
    // (the first layer stages 2 words per pixel and [NB][20] words of weights)
    constexpr int NLDS_IN = (FUSE & kFuseF32In) ? IH * IW * 2 + NB * 20 : IH * IW * 4 + 9 * NB * 4;
    constexpr int NLDS = NLDS_IN > TH * TW * NB / 8 ? NLDS_IN : TH * TW * NB / 8;
    static_assert(TH * TW * NB * 2 <= NLDS * 16, "output tile fits the staging LDS");
    __shared__ uint4 s_lds[NLDS];
    uint4* const s_in = s_lds;
    uint4* const s_wt = s_lds + IH * IW * 4;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, kq = lane >> 4;
    uint32_t bid = blockIdx.x;
    const uint32_t tx = bid % tiles_x;
    bid /= tiles_x;
    const uint32_t ty = bid % tiles_y, img = bid / tiles_y;
    const int H = (int)A.h, W = (int)A.w;
    const int y0 = (int)ty * TH, x0 = (int)tx * TW;
    const uint32_t cbase = blockIdx.y * NB;
    const uint32_t cin = A.c1 + A.c2;
    const uint32_t hh = A.h >> 1, wh = A.w >> 1;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[m][j] = f32x4{0, 0, 0, 0};
    // the lane's output channels' biases, loaded with the first slice (not
    // one load and wait each in the epilogue)
    float bias_r[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) bias_r[j] = A.bias ? A.bias[cbase + j * 16 + r] : 0.0f;
    // the fused head's inputs likewise: this thread's pixel's residual
    // channels and one of the head's 132 weights and biases (to LDS in the
    // epilogue, once the slices' LDS reads are done)
    const bool head = (FUSE & kFuseHead) && NB == 32 && A.head_out;
    float4 hin = make_float4(0, 0, 0, 0);
    float hwv = 0.0f;
    if (head) {
        const int y = y0 + tid / TW, x = x0 + tid % TW;
        if (y < H && x < W) hin = *reinterpret_cast<const float4*>(A.head_in + (((size_t)img * H + y) * W + x) * A.head_cin);
        if (tid < 128) hwv = A.head_w[tid];
        else if (tid < 132) hwv = A.head_b[tid - 128];
    }

    constexpr int NIN = IH * IW * 4, NWT = 9 * NB * 4;
    constexpr int KIN = (NIN + kThreads - 1) / kThreads, KWT = (NWT + kThreads - 1) / kThreads;
    uint4 rin[KIN], rwt[KWT];
    auto fetch = [&](uint32_t c0) {
        const bool from_a = c0 < A.c1;
#pragma unroll
        for (int k = 0; k < KIN; ++k) {
            const int i = tid + k * kThreads;
            const int q = i & 3, pix = i >> 2;
            const int iy = pix / IW, ix = pix - iy * IW;
            const int y = y0 + iy - 1, x = x0 + ix - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (i < NIN && y >= 0 && y < H && x >= 0 && x < W) {
                const uint16_t* src;
                if (!from_a)
                    src = A.b + (((size_t)img * H + y) * W + x) * A.c2 + (c0 - A.c1) + q * 8;
                else if (A.up_a)
                    src = A.a + (((size_t)img * hh + (y >> 1)) * wh + (x >> 1)) * A.c1 + c0 + q * 8;
                else
                    src = A.a + (((size_t)img * H + y) * W + x) * A.c1 + c0 + q * 8;
                v = *reinterpret_cast<const uint4*>(src);
            }
            rin[k] = v;
        }
#pragma unroll
        for (int k = 0; k < KWT; ++k) {
            const int i = tid + k * kThreads;
            const int q = i & 3, rest = i >> 2;
            const int co = rest % NB, t = rest / NB;
            rwt[k] = i < NWT ? *reinterpret_cast<const uint4*>(A.wts + ((size_t)(cbase + co) * 9 + t) * cin + c0 + q * 8)
                             : make_uint4(0, 0, 0, 0);
        }
    };
    if constexpr ((FUSE & kFuseF32In) != 0) {
        // first layer: the fp32 image's in_c <= 16 channels as 16 bf16 (two
        // 16-byte words per pixel) and K = 9 taps x 16 channels, padded to 160:
        // K-step s, lane group kq holds K-group g = 4s + kq = (tap g / 2,
        // channels 8 (g & 1) .. + 7), so 5 K-steps cover what 9 slice-taps of
        // 32 zero-padded channels did.  Weights [cout][160] (k_pack_first).
        static_assert(NT == 2, "the first layer has 32 output channels");
        constexpr int KW = 20;  // 16-byte words of K per output channel
        static_assert(IH * IW * 2 + NB * KW <= NLDS, "the first layer's staging fits");
        uint4* const f_in = s_lds;
        uint4* const f_wt = s_lds + IH * IW * 2;
        // every load of the thread is issued before the first is converted
        // and stored (one memory latency per block, not one per load)
        constexpr int NPX = IH * IW * 2, KPX = (NPX + kThreads - 1) / kThreads;
        constexpr int NFW = NB * KW, KFW = (NFW + kThreads - 1) / kThreads;
        float4 fv[KPX][2];
        uint4 fw[KFW];
#pragma unroll
        for (int k = 0; k < KPX; ++k) {
            const int i = tid + k * kThreads;
            const int h = i & 1, pix = i >> 1;
            const int iy = pix / IW, ix = pix - iy * IW;
            const int y = y0 + iy - 1, x = x0 + ix - 1;
            fv[k][0] = fv[k][1] = make_float4(0, 0, 0, 0);
            if (i < NPX && y >= 0 && y < H && x >= 0 && x < W) {
                const float* px = A.in_f32 + (((size_t)img * H + y) * W + x) * A.in_c;
#pragma unroll
                for (int g = 0; g < 2; ++g) {
                    const uint32_t c = (uint32_t)(8 * h + 4 * g);
                    if (c < A.in_c) fv[k][g] = *reinterpret_cast<const float4*>(px + c);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < KFW; ++k) {
            const int i = tid + k * kThreads;
            fw[k] = i < NFW ? *reinterpret_cast<const uint4*>(A.wts + (size_t)(cbase + i / KW) * (8 * KW) + 8 * (i % KW))
                            : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int k = 0; k < KPX; ++k) {
            const int i = tid + k * kThreads;
            uint4 v;
            uint16_t* v8 = reinterpret_cast<uint16_t*>(&v);
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                v8[4 * g] = f2bf(fv[k][g].x);  // zero (outside the image, or past in_c) stays zero
                v8[4 * g + 1] = f2bf(fv[k][g].y);
                v8[4 * g + 2] = f2bf(fv[k][g].z);
                v8[4 * g + 3] = f2bf(fv[k][g].w);
            }
            if (i < NPX) f_in[i] = v;
        }
#pragma unroll
        for (int k = 0; k < KFW; ++k) {
            const int i = tid + k * kThreads;
            if (i < NFW) f_wt[i] = fw[k];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 5; ++ks) {
            const int g = 4 * ks + kq, t = g >> 1, h = g & 1;
            const int dy = t / 3, dx = t - 3 * (t / 3);
            bf16x8 bw[NT];
#pragma unroll
            for (int j = 0; j < NT; ++j) bw[j] = __builtin_bit_cast(bf16x8, f_wt[(j * 16 + r) * KW + 4 * ks + kq]);
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const int row = wave * (MT / 2) + (m >> 1), col = (m & 1) * 16 + r;
                const bf16x8 af = t < 9 ? __builtin_bit_cast(bf16x8, f_in[2 * ((row + dy) * IW + col + dx) + h]) : zero8();
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bw[j], acc[m][j], 0, 0, 0);
            }
        }
        __syncthreads();  // the LDS reads are done before the epilogue stages the outputs
    } else {
    fetch(0);
    for (uint32_t c0 = 0; c0 < cin; c0 += 32) {
#pragma unroll
        for (int k = 0; k < KIN; ++k) {
            const int i = tid + k * kThreads;
            if (i < NIN) s_in[lds_word(i >> 2, i & 3)] = rin[k];
        }
#pragma unroll
        for (int k = 0; k < KWT; ++k) {
            const int i = tid + k * kThreads;
            if (i < NWT) s_wt[lds_word(i >> 2, i & 3)] = rwt[k];
        }
        __syncthreads();
        if (!(FUSE & kOneSlice) && c0 + 32 < cin) fetch(c0 + 32);
#pragma unroll 3
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3, dx = t - 3 * (t / 3);  // offsets in the haloed tile
            bf16x8 bw[NT];
#pragma unroll
            for (int j = 0; j < NT; ++j) bw[j] = __builtin_bit_cast(bf16x8, s_wt[lds_word(t * NB + j * 16 + r, kq)]);
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const int row = wave * (MT / 2) + (m >> 1), col = (m & 1) * 16 + r;
                const bf16x8 af = __builtin_bit_cast(bf16x8, s_in[lds_word((row + dy) * IW + col + dx, kq)]);
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bw[j], acc[m][j], 0, 0, 0);
            }
        }
        __syncthreads();  // the slice's LDS reads are done before the next is staged
    }
    }
    // D[m][n]: lane holds pixels 4*kq .. 4*kq + 3 of its fragment, output channel r
    if (!A.out_f32 && A.split == A.cout && !A.mask1) {
        // forward pass: bf16 outputs through LDS, then 16-byte stores of 8
        // consecutive channels of a pixel (the LDS reads of the last slice are
        // done: the loop ends with a barrier)
        uint16_t* s_out = reinterpret_cast<uint16_t*>(s_lds);
        // head: its weights and biases go to LDS past the outputs
        float* s_head = reinterpret_cast<float*>(s_lds) + TH * TW * NB / 2;
        static_assert(!(FUSE & kFuseHead) || TH * TW == kThreads, "the head takes one pixel per thread");
        static_assert(TH * TW * NB * 2 + 132 * 4 <= NLDS * 16, "the head's weights fit past the outputs");
        if (head && tid < 132) s_head[tid] = hwv;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int cl = j * 16 + r;
            const float bias = bias_r[j];
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const int row = wave * (MT / 2) + (m >> 1);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float v = acc[m][j][e] + bias;
                    if (A.relu) v = v > 0.0f ? v : 0.0f;
                    s_out[(row * TW + (m & 1) * 16 + kq * 4 + e) * NB + cl] = f2bf(v);
                }
            }
        }
        __syncthreads();
        constexpr int CH = NB / 8;  // 16-byte chunks per pixel
        if (A.out) {
            uint16_t* out = static_cast<uint16_t*>(A.out);
            for (int i = tid; i < TH * TW * CH; i += kThreads) {
                const int pix = i / CH, c = i - pix * CH;
                const int y = y0 + pix / TW, x = x0 + pix % TW;
                if (y >= H || x >= W) continue;
                *reinterpret_cast<uint4*>(out + (((size_t)img * H + y) * W + x) * A.cout + cbase + 8 * c) =
                    *reinterpret_cast<const uint4*>(s_out + pix * NB + 8 * c);
            }
        }
        if ((FUSE & kFusePool) && A.pool) {
            // k_pool's sum: the four bf16 outputs of a 2x2 cell in row-major
            // order, times 0.25 (the tile's origin is even)
            const int Hh = H >> 1, Wh = W >> 1;
            for (int i = tid; i < (TH / 2) * (TW / 2) * CH; i += kThreads) {
                const int pp = i / CH, c = i - pp * CH;
                const int py = pp / (TW / 2), px = pp - py * (TW / 2);
                const int gy = (y0 >> 1) + py, gx = (x0 >> 1) + px;
                if (gy >= Hh || gx >= Wh) continue;
                float sm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int pix = (2 * py + (k >> 1)) * TW + 2 * px + (k & 1);
                    const uint4 v = *reinterpret_cast<const uint4*>(s_out + pix * NB + 8 * c);
                    const uint16_t* v8 = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                    for (int j = 0; j < 8; ++j) sm[j] += bf2f(v8[j]);
                }
                uint4 o;
                uint16_t* o8 = reinterpret_cast<uint16_t*>(&o);
#pragma unroll
                for (int j = 0; j < 8; ++j) o8[j] = f2bf(sm[j] * 0.25f);
                *reinterpret_cast<uint4*>(A.pool + (((size_t)img * Hh + gy) * Wh + gx) * A.cout + cbase + 8 * c) = o;
            }
        }
        if (head) {
            // k_head's sums for this thread's pixel, in its order
            const int pix = tid, y = y0 + pix / TW, x = x0 + pix % TW;
            if (y < H && x < W) {
                const size_t p = ((size_t)img * H + y) * W + x;
                float sm[4] = {hin.x + s_head[128], hin.y + s_head[129], hin.z + s_head[130], hin.w + s_head[131]};
#pragma unroll
                for (int c = 0; c < 32; c += 8) {
                    const uint4 v = *reinterpret_cast<const uint4*>(s_out + pix * NB + c);
                    const uint16_t* v8 = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const float yv = bf2f(v8[j]);
#pragma unroll
                        for (int o = 0; o < 4; ++o) sm[o] += s_head[o * 32 + c + j] * yv;
                    }
                }
                *reinterpret_cast<float4*>(A.head_out + p * 4) = make_float4(sm[0], sm[1], sm[2], sm[3]);
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const uint32_t co = cbase + j * 16 + r;
        const float bias = bias_r[j];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int y = y0 + wave * (MT / 2) + (m >> 1);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int x = x0 + (m & 1) * 16 + kq * 4 + e;
                if (y >= H || x >= W) continue;
                const size_t p = ((size_t)img * H + y) * W + x;
                float v = acc[m][j][e] + bias;
                if (A.relu) v = v > 0.0f ? v : 0.0f;
                if (co < A.split) {
                    const size_t o = p * A.split + co;
                    if (A.mask1 && !bf_pos(A.mask1[o])) v = 0.0f;
                    if (A.out_f32)
                        static_cast<float*>(A.out)[o] = v;
                    else
                        static_cast<uint16_t*>(A.out)[o] = f2bf(v);
                } else {
                    const uint32_t c2o = A.cout - A.split;
                    const size_t o = p * c2o + (co - A.split);
                    if (A.mask2 && !bf_pos(A.mask2[o])) v = 0.0f;
                    uint16_t* o2 = static_cast<uint16_t*>(A.out2);
                    if (A.acc2) v += bf2f(o2[o]);
                    o2[o] = f2bf(v);
                }
            }
        }
    }
}

// Weight (and bias) gradient of k_conv3x3:
//   dw[co][t][ci] += sum_p g[p][co] * x_t[p][ci],  db[co] += sum_p g[p][co]
// x_t is the layer input shifted by tap t, read with the same concat /
// upsample rules as the forward pass.  A block owns a 64 x 64 (co, ci) tile
// of one tap over a range of pixels (one split of K = pixels); 32-pixel
// chunks of g and x are staged in LDS transposed (channel-major) so each
// lane reads its 8 consecutive pixels as one 16-byte operand, and the next
// chunk's loads are in flight while the current chunk's MFMAs run.  Each
// block stores its partial tile (and, for the centre tap of the first input
// tile, its partial bias) in `part`; k_wgrad_reduce adds the splits in order,
// so the result does not depend on the order blocks finish in.
struct WgradArgs {
    const uint16_t* a;
    const uint16_t* b;
    const uint16_t* g;  // [pixels][cout]
    float* dw;          // [cout][9][c1 + c2]
    float* db;          // [cout] or null
    float* part;        // [tiles][splits][64 x 64] partial tiles, then [co_tiles][splits][64] partial biases
    uint32_t c1, c2, cout, n, h, w;
    int up_a;
    uint32_t ci_tiles, co_tiles, chunk_per_block, splits;  // chunks of 32 pixels per block, blocks per tile
};

constexpr int kWChunk = 32;
constexpr int kWPad = kWChunk + 8;
constexpr int kWTile = 64 * 64;

__global__ __launch_bounds__(kThreads) void k_wgrad(WgradArgs A)
{
    __shared__ uint16_t sg[64 * kWPad];
    __shared__ uint16_t sx[64 * kWPad];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, kq = lane >> 4;
    uint32_t bid = blockIdx.y;
    const uint32_t ct = bid % A.ci_tiles;
    bid /= A.ci_tiles;
    const uint32_t ot = bid % A.co_tiles;
    const int t = (int)(bid / A.co_tiles);
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const uint32_t cin = A.c1 + A.c2;
    const uint32_t ci0 = ct * 64, co0 = ot * 64;
    const uint32_t hw = A.h * A.w, P = A.n * hw;
    const uint32_t hh = A.h >> 1, wh = A.w >> 1;
    const uint32_t p_begin = blockIdx.x * A.chunk_per_block * kWChunk;
    const uint32_t p_end = min(P, p_begin + A.chunk_per_block * kWChunk);
    const bool do_bias = A.db && t == 4 && ct == 0;

    // loader role: pixel lp of the chunk, channels lc .. lc + 7 of the tile
    const int lp = tid >> 3, lc = (tid & 7) * 8;
    // compute role: wave -> (co half, ci half) of the tile
    const int wco = (wave >> 1) * 32, wci = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
    float bsum = 0.0f;

    auto load = [&](uint32_t p0, uint4& gv, uint4& xv) {
        const uint32_t p = p0 + lp;
        gv = make_uint4(0, 0, 0, 0);
        xv = make_uint4(0, 0, 0, 0);
        if (p >= p_end) return;
        const uint32_t co = co0 + lc;
        if (co < A.cout) gv = *reinterpret_cast<const uint4*>(A.g + (size_t)p * A.cout + co);
        const uint32_t ci = ci0 + lc;
        const uint32_t pn = p / hw, rem = p - pn * hw;
        const int y = (int)(rem / A.w) + dy, x = (int)(rem % A.w) + dx;
        if (ci < cin && y >= 0 && y < (int)A.h && x >= 0 && x < (int)A.w) {
            const uint16_t* src;
            if (ci < A.c1)
                src = A.up_a ? A.a + (((size_t)pn * hh + (y >> 1)) * wh + (x >> 1)) * A.c1 + ci
                             : A.a + (((size_t)pn * A.h + y) * A.w + x) * A.c1 + ci;
            else
                src = A.b + (((size_t)pn * A.h + y) * A.w + x) * A.c2 + (ci - A.c1);
            xv = *reinterpret_cast<const uint4*>(src);
        }
    };
    uint4 gv, xv;
    load(p_begin, gv, xv);
    for (uint32_t p0 = p_begin; p0 < p_end; p0 += kWChunk) {
        __syncthreads();  // the previous chunk's reads are done
        const uint16_t* g8 = reinterpret_cast<const uint16_t*>(&gv);
        const uint16_t* x8 = reinterpret_cast<const uint16_t*>(&xv);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            sg[(lc + j) * kWPad + lp] = g8[j];
            sx[(lc + j) * kWPad + lp] = x8[j];
        }
        if (p0 + kWChunk < p_end) load(p0 + kWChunk, gv, xv);
        __syncthreads();
        if (do_bias && tid < 64) {
            for (int q = 0; q < kWChunk; ++q) bsum += bf2f(sg[tid * kWPad + q]);
        }
        bf16x8 af[2], bx[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) af[i] = load8(&sg[(wco + i * 16 + r) * kWPad + kq * 8]);
#pragma unroll
        for (int j = 0; j < 2; ++j) bx[j] = load8(&sx[(wci + j * 16 + r) * kWPad + kq * 8]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bx[j], acc[i][j], 0, 0, 0);
    }
    float* tile = A.part + ((size_t)blockIdx.y * A.splits + blockIdx.x) * kWTile;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                tile[(wco + i * 16 + kq * 4 + e) * 64 + wci + j * 16 + r] = acc[i][j][e];
    if (do_bias && tid < 64) {
        float* bp = A.part + (size_t)gridDim.y * A.splits * kWTile;
        bp[((size_t)ot * A.splits + blockIdx.x) * 64 + tid] = bsum;
    }
}

// Second pass of k_wgrad: dw[co][t][ci] += the splits of its tile in split
// order (and db[co] likewise), one thread per weight.
__global__ void k_wgrad_reduce(WgradArgs A)
{
    const uint32_t cin = A.c1 + A.c2;
    const size_t nw = (size_t)A.cout * 9 * cin;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nw) {
        const uint32_t ci = (uint32_t)(i % cin), rest = (uint32_t)(i / cin);
        const uint32_t t = rest % 9, co = rest / 9;
        const uint32_t tile = (t * A.co_tiles + co / 64) * A.ci_tiles + ci / 64;
        const float* p = A.part + (size_t)tile * A.splits * kWTile + (co % 64) * 64 + ci % 64;
        float s = 0.0f;
        for (uint32_t k = 0; k < A.splits; ++k) s += p[(size_t)k * kWTile];
        A.dw[i] += s;
    } else if (A.db && i < nw + A.cout) {
        const uint32_t co = (uint32_t)(i - nw);
        const float* p = A.part + (size_t)9 * A.co_tiles * A.ci_tiles * A.splits * kWTile +
                         (size_t)(co / 64) * A.splits * 64 + co % 64;
        float s = 0.0f;
        for (uint32_t k = 0; k < A.splits; ++k) s += p[(size_t)k * 64];
        A.db[co] += s;
    }
}

// 2x2 average pooling, 8 channels per thread: out[n][y][x] = mean of the
// four input pixels (h, w even).
__global__ void k_pool(const uint16_t* in, uint16_t* out, uint32_t n, uint32_t h, uint32_t w, uint32_t c)
{
    const uint32_t ho = h >> 1, wo = w >> 1, c8 = c >> 3;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * ho * wo * c8) return;
    const uint32_t cc = (uint32_t)(i % c8) * 8;
    size_t q = i / c8;
    const uint32_t x = q % wo;
    q /= wo;
    const uint32_t y = q % ho;
    const uint32_t b = (uint32_t)(q / ho);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        const size_t src = (((size_t)b * h + 2 * y + (k >> 1)) * w + 2 * x + (k & 1)) * c + cc;
        const uint4 v = *reinterpret_cast<const uint4*>(in + src);
        const uint16_t* v8 = reinterpret_cast<const uint16_t*>(&v);
        for (int j = 0; j < 8; ++j) s[j] += bf2f(v8[j]);
    }
    uint4 o;
    uint16_t* o8 = reinterpret_cast<uint16_t*>(&o);
    for (int j = 0; j < 8; ++j) o8[j] = f2bf(s[j] * 0.25f);
    *reinterpret_cast<uint4*>(out + (((size_t)b * ho + y) * wo + x) * c + cc) = o;
}

// Backward of k_pool into the gradient of its input, which is the ReLU
// output `act` of a convolution: gin += 0.25 * gout (upsampled), where act > 0.
__global__ void k_pool_back(const uint16_t* gout, const uint16_t* act, uint16_t* gin, uint32_t n, uint32_t h,
                            uint32_t w, uint32_t c)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c8 = c >> 3;
    if (i >= (size_t)n * h * w * c8) return;
    const uint32_t cc = (uint32_t)(i % c8) * 8;
    size_t q = i / c8;
    const uint32_t x = q % w;
    q /= w;
    const uint32_t y = q % h;
    const uint32_t b = (uint32_t)(q / h);
    const size_t o = (((size_t)b * h + y) * w + x) * c + cc;
    const size_t s = (((size_t)b * (h >> 1) + (y >> 1)) * (w >> 1) + (x >> 1)) * c + cc;
    const uint4 gv = *reinterpret_cast<const uint4*>(gout + s);
    const uint4 av = *reinterpret_cast<const uint4*>(act + o);
    uint4 iv = *reinterpret_cast<const uint4*>(gin + o);
    const uint16_t* g8 = reinterpret_cast<const uint16_t*>(&gv);
    const uint16_t* a8 = reinterpret_cast<const uint16_t*>(&av);
    uint16_t* i8 = reinterpret_cast<uint16_t*>(&iv);
    for (int j = 0; j < 8; ++j)
        if (bf_pos(a8[j])) i8[j] = f2bf(bf2f(i8[j]) + 0.25f * bf2f(g8[j]));
    *reinterpret_cast<uint4*>(gin + o) = iv;
}

// Backward of the nearest upsampling folded into k_conv3x3's `a` input:
// gsmall[n][y][x] = sum of the 2x2 full-resolution gradients, where the
// half-resolution activation act > 0.  (h, w: the half resolution.)
__global__ void k_up_back(const uint16_t* gbig, const uint16_t* act, uint16_t* gsmall, uint32_t n, uint32_t h,
                          uint32_t w, uint32_t c)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c8 = c >> 3;
    if (i >= (size_t)n * h * w * c8) return;
    const uint32_t cc = (uint32_t)(i % c8) * 8;
    size_t q = i / c8;
    const uint32_t x = q % w;
    q /= w;
    const uint32_t y = q % h;
    const uint32_t b = (uint32_t)(q / h);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        const size_t src = (((size_t)b * 2 * h + 2 * y + (k >> 1)) * 2 * w + 2 * x + (k & 1)) * c + cc;
        const uint4 v = *reinterpret_cast<const uint4*>(gbig + src);
        const uint16_t* v8 = reinterpret_cast<const uint16_t*>(&v);
        for (int j = 0; j < 8; ++j) s[j] += bf2f(v8[j]);
    }
    const size_t o = (((size_t)b * h + y) * w + x) * c + cc;
    const uint4 av = *reinterpret_cast<const uint4*>(act + o);
    const uint16_t* a8 = reinterpret_cast<const uint16_t*>(&av);
    uint4 ov;
    uint16_t* o8 = reinterpret_cast<uint16_t*>(&ov);
    for (int j = 0; j < 8; ++j) o8[j] = bf_pos(a8[j]) ? f2bf(s[j]) : (uint16_t)0;
    *reinterpret_cast<uint4*>(gsmall + o) = ov;
}

// Network input: fp32 [pixels][cin] -> bf16 [pixels][32], channels past cin zero.
__global__ void k_pack_input(const float* in, uint16_t* out, size_t pixels, uint32_t cin)
{
    // one thread per 8 channels of a pixel: one 16-byte store
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels * 4) return;
    const uint32_t c0 = (uint32_t)(i & 3) * 8;
    const size_t p = i >> 2;
    uint4 o;
    uint16_t* o8 = reinterpret_cast<uint16_t*>(&o);
#pragma unroll
    for (int j = 0; j < 8; ++j) o8[j] = c0 + j < cin ? f2bf(in[p * cin + c0 + j]) : (uint16_t)0;
    *reinterpret_cast<uint4*>(out + p * 32 + c0) = o;
}

// 1x1 head: out[p][o] = res[p][o] + hb[o] + sum_c hw[o][c] * y[p][c]
// (o < 4, c < 32; res = the noisy AO and bent normal, input channels 0..3).
constexpr int kHeadIn = 32;
constexpr int kOut = 4;

__global__ void k_head(const uint16_t* y, const float* in, uint32_t cin, const float* hw, const float* hb, float* out,
                       size_t pixels)
{
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    float s[kOut];
    for (int o = 0; o < kOut; ++o) s[o] = in[p * cin + o] + hb[o];
    for (int c = 0; c < kHeadIn; c += 8) {
        const uint4 v = *reinterpret_cast<const uint4*>(y + p * kHeadIn + c);
        const uint16_t* v8 = reinterpret_cast<const uint16_t*>(&v);
        for (int j = 0; j < 8; ++j) {
            const float yv = bf2f(v8[j]);
            for (int o = 0; o < kOut; ++o) s[o] += hw[o * kHeadIn + c + j] * yv;
        }
    }
    for (int o = 0; o < kOut; ++o) out[p * kOut + o] = s[o];
}

// Masked L1 loss and the backward of the head.  valid = input channel
// `mask_ch` > 0.5; L = (1/n_valid) sum_valid sum_o |out - target|.
// Writes gy = dL/dy (bf16, zero where y <= 0) and accumulates the head's
// weight and bias gradients and the loss: grid-stride over pixels, sums kept
// per lane, reduced per wave, one fp32 atomic per wave and value.
__device__ inline float wave_sum(float v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__global__ __launch_bounds__(kThreads) void k_head_back(const uint16_t* y, const float* in, uint32_t cin,
                                                        uint32_t mask_ch, const float* out, const float* target,
                                                        const float* hw, float inv_valid, uint16_t* gy,
                                                        float* ghw, float* ghb, float* loss, size_t pixels)
{
    float gw[kOut][kHeadIn], gb[kOut], l = 0.0f;
    for (int o = 0; o < kOut; ++o) {
        gb[o] = 0.0f;
        for (int c = 0; c < kHeadIn; ++c) gw[o][c] = 0.0f;
    }
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += (size_t)gridDim.x * blockDim.x) {
        const bool valid = in[p * cin + mask_ch] > 0.5f;
        float d[kOut];
        for (int o = 0; o < kOut; ++o) {
            const float e = out[p * kOut + o] - target[p * kOut + o];
            d[o] = 0.0f;
            if (valid) {
                l += fabsf(e);
                d[o] = e > 0.0f ? inv_valid : (e < 0.0f ? -inv_valid : 0.0f);
            }
            gb[o] += d[o];
        }
        for (int c = 0; c < kHeadIn; ++c) {
            const uint16_t u = y[p * kHeadIn + c];
            const float yv = bf2f(u);
            float gsum = 0.0f;
            for (int o = 0; o < kOut; ++o) {
                gsum += d[o] * hw[o * kHeadIn + c];
                gw[o][c] += d[o] * yv;
            }
            gy[p * kHeadIn + c] = bf_pos(u) ? f2bf(gsum) : (uint16_t)0;
        }
    }
    const bool lead = (threadIdx.x & 63) == 0;
    const float lt = wave_sum(l);
    if (lead) atomicAdd(loss, lt * inv_valid);
    for (int o = 0; o < kOut; ++o) {
        const float bt = wave_sum(gb[o]);
        if (lead) atomicAdd(&ghb[o], bt);
        for (int c = 0; c < kHeadIn; ++c) {
            const float wt = wave_sum(gw[o][c]);
            if (lead) atomicAdd(&ghw[o * kHeadIn + c], wt);
        }
    }
}

// Number of masked pixels (input channel mask_ch > 0.5).
__global__ void k_count_valid(const float* in, uint32_t cin, uint32_t mask_ch, size_t pixels, uint32_t* count)
{
    uint32_t c = 0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += (size_t)gridDim.x * blockDim.x)
        c += in[p * cin + mask_ch] > 0.5f;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(count, c);
}

// Adam on fp32 master weights; refreshes the bf16 copy.
__global__ void k_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, size_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

// fp32 -> bf16 copy of weights
__global__ void k_to_bf16(const float* w, uint16_t* wb, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(w[i]);
}

// Weights of the input-gradient convolution: wt[ci][8 - t][co] = w[co][t][ci].
__global__ void k_flip_weights(const uint16_t* w, uint16_t* wt, uint32_t cout, uint32_t cin)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)cout * 9 * cin) return;
    const uint32_t ci = (uint32_t)(i % cin);
    const uint32_t t = (uint32_t)((i / cin) % 9);
    const uint32_t co = (uint32_t)(i / ((size_t)cin * 9));
    wt[((size_t)ci * 9 + (8 - t)) * cout + co] = w[i];
}

// The first layer's weights [cout][9][cin] (bf16) as [cout][160]: K index
// 16 t + c holds tap t, channel c < 16; the rest is zero (k_conv3x3 with
// kFuseF32In).
__global__ void k_pack_first(const uint16_t* w, uint16_t* wk, uint32_t cout, uint32_t cin)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)cout * 160) return;
    const uint32_t k = (uint32_t)(i % 160), co = (uint32_t)(i / 160);
    const uint32_t t = k >> 4, c = k & 15;
    wk[i] = (t < 9 && c < cin) ? w[((size_t)co * 9 + t) * cin + c] : (uint16_t)0;
}

__global__ void k_zero(float* p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_zero_u32(uint32_t* p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace dn
}  // namespace fornos
