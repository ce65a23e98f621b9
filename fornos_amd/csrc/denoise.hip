// denoise.hip -- the U-Net denoiser behind include/fornos_denoise.h.
//
// Layers (L levels, C base channels, C_l = C * 2^l):
//   encoder level l:  E[l][0] = relu(conv(l ? pool(E[l-1][1]) : x))   C_l
//                     E[l][1] = relu(conv(E[l][0]))                     C_l
//   decoder level l = L-2 .. 0, "below" = E[L-1][1] or D[l+1][1]:
//                     D[l][0] = relu(conv(concat(up(below), E[l][1])))  C_l
//                     D[l][1] = relu(conv(D[l][0]))                     C_l
//   head:             out = x[0..3] + hb + hw * D[0][1]   (1x1, residual)
// Activations and activation gradients are bf16 NHWC; weights are fp32
// masters with bf16 copies; every convolution accumulates in fp32 on the
// matrix cores.  x is packed to 32 bf16 channels (12 used).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/fornos_denoise.h"
#include "dn_kernels.h"
#include "hiphost.h"

using namespace fornos;
using namespace fornos::dn;

namespace {

static_assert(FORNOS_DN_IN <= 16 && FORNOS_DN_IN % 4 == 0, "the first layer reads at most 16 fp32 channels, 4 at a time");

struct Layer {
    uint32_t c1, c2, cout, up, level;
    uint64_t w_off, b_off;  // in the parameter vector
    uint64_t f_off;         // in the flipped-weight buffer
};

// Tiles of 8 x 32 output pixels, for 64 output channels (or 32 when cout is
// not a multiple of 64) per block.  (Tiles of 16 x 32 pixels for 32 output
// channels measured no faster.)
void launch_conv(const ConvArgs& A, hipStream_t st)
{
    constexpr int MT = 4, TH = 2 * MT, TW = 32;
    const uint32_t tx = div_up(A.w, TW), ty = div_up(A.h, TH);
    const uint32_t tiles = A.n * tx * ty;
    const dim3 grid32(tiles, A.cout / 32), grid64(tiles, A.cout / 64);
    const bool one = A.c1 + A.c2 == 32;  // one input slice: no next-slice prefetch registers
    if (A.in_f32)  // the first layer (32 output channels)
        hipLaunchKernelGGL((k_conv3x3<MT, 2, kFuseF32In>), grid32, dim3(kThreads), 0, st, A, tx, ty);
    else if (A.head_out)  // the last layer (32 output channels)
        if (one)
            hipLaunchKernelGGL((k_conv3x3<MT, 2, kFuseHead | kOneSlice>), grid32, dim3(kThreads), 0, st, A, tx, ty);
        else
            hipLaunchKernelGGL((k_conv3x3<MT, 2, kFuseHead>), grid32, dim3(kThreads), 0, st, A, tx, ty);
    else if (A.cout % 64 == 0)
        if (A.pool)
            hipLaunchKernelGGL((k_conv3x3<MT, 4, kFusePool>), grid64, dim3(kThreads), 0, st, A, tx, ty);
        else if (one)
            hipLaunchKernelGGL((k_conv3x3<MT, 4, kOneSlice>), grid64, dim3(kThreads), 0, st, A, tx, ty);
        else
            hipLaunchKernelGGL((k_conv3x3<MT, 4>), grid64, dim3(kThreads), 0, st, A, tx, ty);
    else if (A.pool)
        if (one)
            hipLaunchKernelGGL((k_conv3x3<MT, 2, kFusePool | kOneSlice>), grid32, dim3(kThreads), 0, st, A, tx, ty);
        else
            hipLaunchKernelGGL((k_conv3x3<MT, 2, kFusePool>), grid32, dim3(kThreads), 0, st, A, tx, ty);
    else if (one)
        hipLaunchKernelGGL((k_conv3x3<MT, 2, kOneSlice>), grid32, dim3(kThreads), 0, st, A, tx, ty);
    else
        hipLaunchKernelGGL((k_conv3x3<MT, 2>), grid32, dim3(kThreads), 0, st, A, tx, ty);
}

// k_wgrad + k_wgrad_reduce; `part` is their partial-tile buffer (grown when
// a layer needs more, after the stream's earlier work).
fornos_status launch_wgrad(WgradArgs A, hipStream_t st, float*& part, size_t& part_cap)
{
    const uint64_t P = (uint64_t)A.n * A.h * A.w;
    const uint32_t cin = A.c1 + A.c2;
    A.ci_tiles = div_up(cin, 64);
    A.co_tiles = div_up(A.cout, 64);
    const uint32_t tiles = 9 * A.ci_tiles * A.co_tiles;
    const uint32_t chunks = div_up(P, kWChunk);
    const uint32_t splits = std::max<uint32_t>(1, std::min<uint32_t>(chunks, div_up(4096, tiles)));
    A.chunk_per_block = div_up(chunks, splits);
    A.splits = div_up(chunks, A.chunk_per_block);
    const size_t need = (size_t)A.splits * (tiles * (size_t)kWTile + A.co_tiles * (size_t)64);
    if (need > part_cap) {
        HIP_TRY(hipStreamSynchronize(st));
        dev_free(part);
        part = nullptr;
        part_cap = 0;
        if (fornos_status e = dev_alloc(&part, need)) return e;
        part_cap = need;
    }
    A.part = part;
    hipLaunchKernelGGL(k_wgrad, dim3(A.splits, tiles), dim3(kThreads), 0, st, A);
    const size_t nred = (size_t)A.cout * 9 * cin + (A.db ? A.cout : 0);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(div_up(nred, kThreads)), dim3(kThreads), 0, st, A);
    HIP_TRY(hipGetLastError());
    return FORNOS_OK;
}

template <typename K, typename... Args>
void launch_1d(K kernel, size_t n, hipStream_t st, Args... args)
{
    if (n == 0) return;
    hipLaunchKernelGGL(kernel, dim3(div_up(n, kThreads)), dim3(kThreads), 0, st, args...);
}

bool conv_shape_ok(uint32_t c1, uint32_t c2, uint32_t cout)
{
    return c1 % 32 == 0 && c2 % 32 == 0 && c1 + c2 > 0 && cout % 32 == 0 && cout > 0;
}

}  // namespace

struct fornos_denoiser {
    uint32_t levels = 4, base = 32;
    std::vector<Layer> layers;
    uint64_t n_params = 0, n_flip = 0, head_w = 0, head_b = 0;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wb = nullptr, *wflip = nullptr;
    uint16_t* wfirst = nullptr;  // the first layer's weights packed for the fp32-input kernel
    uint64_t step = 0;
    uint64_t cap_pixels = 0;
    uint16_t* x0 = nullptr;
    std::vector<uint16_t*> act, gact, pool, gpool, gup;
    float* out = nullptr;
    uint32_t* valid = nullptr;
    float* loss = nullptr;
    float* wpart = nullptr;  // weight-gradient partial tiles (launch_wgrad)
    size_t wpart_cap = 0;
    hipStream_t st = nullptr;  // the default stream

    ~fornos_denoiser()
    {
        (void)hipDeviceSynchronize();
        for (void* p : {(void*)params, (void*)grads, (void*)m, (void*)v, (void*)wb, (void*)wflip, (void*)wfirst,
                        (void*)wpart})
            dev_free(p);
        free_acts();
    }
    void free_acts()
    {
        dev_free(x0);
        x0 = nullptr;
        for (auto* vec : {&act, &gact, &pool, &gpool, &gup}) {
            for (uint16_t* p : *vec) dev_free(p);
            vec->clear();
        }
        dev_free(out);
        dev_free(valid);
        dev_free(loss);
        out = nullptr;
        valid = nullptr;
        loss = nullptr;
        cap_pixels = 0;
    }
    uint32_t ch(uint32_t l) const { return base << l; }
    uint32_t enc(uint32_t l, uint32_t k) const { return 2 * l + k; }
    uint32_t dec(uint32_t l, uint32_t k) const { return 2 * levels + 2 * (levels - 2 - l) + k; }

    void build_layers()
    {
        layers.clear();
        uint64_t off = 0, foff = 0;
        auto add = [&](uint32_t c1, uint32_t c2, uint32_t cout, uint32_t up, uint32_t level) {
            Layer L{c1, c2, cout, up, level, off, 0, foff};
            off += (uint64_t)cout * 9 * (c1 + c2);
            L.b_off = off;
            off += cout;
            foff += (uint64_t)cout * 9 * (c1 + c2);
            layers.push_back(L);
        };
        for (uint32_t l = 0; l < levels; ++l) {
            add(l == 0 ? 32 : ch(l - 1), 0, ch(l), 0, l);
            add(ch(l), 0, ch(l), 0, l);
        }
        for (int l = (int)levels - 2; l >= 0; --l) {
            add(ch(l + 1), ch(l), ch(l), 1, l);
            add(ch(l), 0, ch(l), 0, l);
        }
        head_w = off;
        off += (uint64_t)kOut * base;
        head_b = off;
        off += kOut;
        n_params = off;
        n_flip = foff;
    }

    fornos_status ensure(uint32_t n, uint32_t h, uint32_t w)
    {
        const uint64_t P = (uint64_t)n * h * w;
        if (P <= cap_pixels) return FORNOS_OK;
        free_acts();
        auto px = [&](uint32_t l) { return P >> (2 * l); };
        if (fornos_status s = dev_alloc(&x0, P * 32)) return s;
        act.assign(layers.size(), nullptr);
        gact.assign(layers.size(), nullptr);
        for (size_t i = 0; i < layers.size(); ++i) {
            const Layer& L = layers[i];
            if (fornos_status s = dev_alloc(&act[i], px(L.level) * L.cout)) return s;
            if (fornos_status s = dev_alloc(&gact[i], px(L.level) * L.cout)) return s;
        }
        pool.assign(levels, nullptr);
        gpool.assign(levels, nullptr);
        gup.assign(levels, nullptr);
        for (uint32_t l = 1; l < levels; ++l) {
            if (fornos_status s = dev_alloc(&pool[l], px(l) * ch(l - 1))) return s;
            if (fornos_status s = dev_alloc(&gpool[l], px(l) * ch(l - 1))) return s;
        }
        for (uint32_t l = 0; l + 1 < levels; ++l)
            if (fornos_status s = dev_alloc(&gup[l], px(l) * ch(l + 1))) return s;
        if (fornos_status s = dev_alloc(&out, P * kOut)) return s;
        if (fornos_status s = dev_alloc(&valid, 1)) return s;
        if (fornos_status s = dev_alloc(&loss, 1)) return s;
        cap_pixels = P;
        return FORNOS_OK;
    }

    // the forward inputs of layer i
    void inputs(uint32_t i, const uint16_t** a, const uint16_t** b) const
    {
        *b = nullptr;
        if (i < 2 * levels) {
            const uint32_t l = i / 2;
            *a = (i & 1) ? act[i - 1] : (l == 0 ? x0 : pool[l]);
        } else {
            const uint32_t j = i - 2 * levels;
            const uint32_t l = levels - 2 - j / 2;
            if (j & 1) {
                *a = act[i - 1];
            } else {
                *a = (l == levels - 2) ? act[enc(levels - 1, 1)] : act[dec(l + 1, 1)];
                *b = act[enc(l, 1)];
            }
        }
    }

    void conv_forward(uint32_t i, uint32_t n, uint32_t h, uint32_t w, const float* in, float* o, bool train)
    {
        const Layer& L = layers[i];
        ConvArgs A{};
        inputs(i, &A.a, &A.b);
        A.wts = wb + L.w_off;
        A.bias = params + L.b_off;
        A.out = act[i];
        A.c1 = L.c1;
        A.c2 = L.c2;
        A.cout = L.cout;
        A.split = L.cout;
        A.n = n;
        A.h = h >> L.level;
        A.w = w >> L.level;
        A.up_a = (int)L.up;
        A.relu = 1;
        if (i == 0 && !train) {
            // the network input straight from the fp32 image, with the
            // weights packed for it (training keeps the packed copy of the
            // image: the weight gradient of this layer reads it)
            A.in_f32 = in;
            A.in_c = FORNOS_DN_IN;
            A.wts = wfirst;
        }
        if (i < 2 * levels && (i & 1) && i / 2 + 1 < levels) A.pool = pool[i / 2 + 1];  // input of the next level
        if (i == dec(0, 1)) {
            A.head_in = in;
            A.head_cin = FORNOS_DN_IN;
            A.head_w = params + head_w;
            A.head_b = params + head_b;
            A.head_out = o;
            if (!train) A.out = nullptr;  // the head's backward reads these in training
        }
        launch_conv(A, st);
    }

    // train: keep what the backward pass reads (the packed input and the
    // last layer's activations); inference reads the fp32 image in the
    // first layer and stores only the head's output from the last.
    void forward(const float* in, uint32_t n, uint32_t h, uint32_t w, float* o, bool train)
    {
        const uint64_t P = (uint64_t)n * h * w;
        if (train)
            launch_1d(k_pack_input, P * 4, st, in, x0, (size_t)P, (uint32_t)FORNOS_DN_IN);
        else  // from the current bf16 weights (training updates them)
            launch_1d(k_pack_first, (size_t)layers[0].cout * 160, st, (const uint16_t*)(wb + layers[0].w_off), wfirst,
                      layers[0].cout, layers[0].c1);
        for (uint32_t l = 0; l < levels; ++l) {
            conv_forward(enc(l, 0), n, h, w, in, o, train);
            conv_forward(enc(l, 1), n, h, w, in, o, train);
        }
        for (int l = (int)levels - 2; l >= 0; --l) {
            conv_forward(dec(l, 0), n, h, w, in, o, train);
            conv_forward(dec(l, 1), n, h, w, in, o, train);
        }
    }

    // weight gradient of layer i, then the gradient of its inputs
    fornos_status conv_backward(uint32_t i, uint32_t n, uint32_t h, uint32_t w)
    {
        const Layer& L = layers[i];
        const uint32_t hl = h >> L.level, wl = w >> L.level;
        WgradArgs W{};
        inputs(i, &W.a, &W.b);
        W.g = gact[i];
        W.dw = grads + L.w_off;
        W.db = grads + L.b_off;
        W.c1 = L.c1;
        W.c2 = L.c2;
        W.cout = L.cout;
        W.n = n;
        W.h = hl;
        W.w = wl;
        W.up_a = (int)L.up;
        if (fornos_status e = launch_wgrad(W, st, wpart, wpart_cap)) return e;
        if (i == 0) return FORNOS_OK;  // no gradient for the network input

        ConvArgs A{};
        A.a = gact[i];
        A.c1 = L.cout;
        A.wts = wflip + L.f_off;
        A.cout = L.c1 + L.c2;
        A.split = L.c1;
        A.n = n;
        A.h = hl;
        A.w = wl;
        const uint64_t P = (uint64_t)n * hl * wl;
        if (i < 2 * levels) {
            const uint32_t l = i / 2;
            if (i & 1) {  // input E[l][0] (ReLU output)
                A.out = gact[i - 1];
                A.mask1 = act[i - 1];
                launch_conv(A, st);
            } else {  // input pool(E[l-1][1]): no ReLU between; pool backward masks
                A.out = gpool[l];
                launch_conv(A, st);
                launch_1d(k_pool_back, (P << 2) * ch(l - 1) / 8, st, (const uint16_t*)gpool[l],
                          (const uint16_t*)act[enc(l - 1, 1)], gact[enc(l - 1, 1)], n, hl << 1, wl << 1, ch(l - 1));
            }
        } else {
            const uint32_t j = i - 2 * levels;
            const uint32_t l = levels - 2 - j / 2;
            if (j & 1) {  // input D[l][0]
                A.out = gact[i - 1];
                A.mask1 = act[i - 1];
                launch_conv(A, st);
            } else {  // inputs up(below) and E[l][1] (its first gradient contribution)
                A.out = gup[l];
                A.out2 = gact[enc(l, 1)];
                A.mask2 = act[enc(l, 1)];
                launch_conv(A, st);
                const uint32_t bl = (l == levels - 2) ? enc(levels - 1, 1) : dec(l + 1, 1);
                launch_1d(k_up_back, (P >> 2) * ch(l + 1) / 8, st, (const uint16_t*)gup[l],
                          (const uint16_t*)act[bl], gact[bl], n, hl >> 1, wl >> 1, ch(l + 1));
            }
        }
        return FORNOS_OK;
    }

    void refresh_weights()
    {
        launch_1d(k_to_bf16, n_params, st, (const float*)params, wb, (size_t)n_params);
        for (const Layer& L : layers)
            launch_1d(k_flip_weights, (size_t)L.cout * 9 * (L.c1 + L.c2), st, (const uint16_t*)(wb + L.w_off),
                      wflip + L.f_off, L.cout, L.c1 + L.c2);
    }
};

namespace {

fornos_status check_image(const fornos_denoiser* d, uint32_t n, uint32_t h, uint32_t w)
{
    const uint32_t q = 1u << (d->levels - 1);
    if (n == 0 || h == 0 || w == 0 || h % q || w % q || h % 2 || w % 2)
        return set_error(FORNOS_INVALID_ARG, "denoiser: image sides must be positive multiples of " +
                                                 std::to_string(std::max(2u, q)));
    if ((uint64_t)n * h * w >= (1ull << 31))
        return set_error(FORNOS_UNSUPPORTED, "denoiser: at most 2^31 pixels per call");
    return FORNOS_OK;
}

}  // namespace

extern "C" {

fornos_status fornos_denoiser_create(uint32_t levels, uint32_t base_channels, uint64_t seed, fornos_denoiser** out)
{
    if (!out) return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_create: null output");
    *out = nullptr;
    if (levels < 2 || levels > 6 || base_channels == 0 || base_channels % 32 || base_channels > 256)
        return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_create: levels in [2,6], base_channels a multiple of 32 up to 256");
    if (base_channels != (uint32_t)kHeadIn)
        return set_error(FORNOS_UNSUPPORTED, "fornos_denoiser_create: the head takes 32 channels");
    if (fornos_status s = require_device()) return s;
    std::unique_ptr<fornos_denoiser> d(new fornos_denoiser());
    d->levels = levels;
    d->base = base_channels;
    d->build_layers();
    for (const Layer& L : d->layers)
        if (!conv_shape_ok(L.c1, L.c2, L.cout)) return set_error(FORNOS_INVALID_ARG, "denoiser: bad layer shape");
    // He-normal weights (fan-in of the real input channels), zero biases, a
    // small head so that the first output is close to the noisy input
    std::vector<float> h(d->n_params, 0.0f);
    std::mt19937_64 rng(seed);
    for (size_t i = 0; i < d->layers.size(); ++i) {
        const Layer& L = d->layers[i];
        const uint32_t cin = L.c1 + L.c2, real = i == 0 ? FORNOS_DN_IN : cin;
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (9.0f * real)));
        for (uint32_t co = 0; co < L.cout; ++co)
            for (uint32_t t = 0; t < 9; ++t)
                for (uint32_t ci = 0; ci < cin; ++ci)
                    h[L.w_off + ((uint64_t)co * 9 + t) * cin + ci] = ci < real ? nd(rng) : 0.0f;
    }
    std::normal_distribution<float> hn(0.0f, 1e-3f);
    for (uint64_t i = 0; i < (uint64_t)kOut * d->base; ++i) h[d->head_w + i] = hn(rng);
    for (float** p : {&d->params, &d->grads, &d->m, &d->v})
        if (fornos_status s = dev_alloc(p, d->n_params)) return s;
    if (fornos_status s = dev_alloc(&d->wb, d->n_params)) return s;
    if (fornos_status s = dev_alloc(&d->wflip, d->n_flip)) return s;
    if (fornos_status s = dev_alloc(&d->wfirst, (size_t)d->layers[0].cout * 160)) return s;
    HIP_TRY(hipMemcpy(d->params, h.data(), sizeof(float) * d->n_params, hipMemcpyHostToDevice));
    launch_1d(k_zero, d->n_params, d->st, d->grads, (size_t)d->n_params);
    launch_1d(k_zero, d->n_params, d->st, d->m, (size_t)d->n_params);
    launch_1d(k_zero, d->n_params, d->st, d->v, (size_t)d->n_params);
    d->refresh_weights();
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    *out = d.release();
    return FORNOS_OK;
}

fornos_status fornos_denoiser_param_count(const fornos_denoiser* d, uint64_t* count)
{
    if (!d || !count) return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_param_count: null argument");
    *count = d->n_params;
    return FORNOS_OK;
}

fornos_status fornos_denoiser_layer_count(const fornos_denoiser* d, uint32_t* count)
{
    if (!d || !count) return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_layer_count: null argument");
    *count = (uint32_t)d->layers.size();
    return FORNOS_OK;
}

fornos_status fornos_denoiser_layer(const fornos_denoiser* d, uint32_t i, uint32_t* c1, uint32_t* c2, uint32_t* cout,
                                    uint32_t* up, uint64_t* offset)
{
    if (!d || i >= d->layers.size()) return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_layer: bad layer index");
    const Layer& L = d->layers[i];
    if (c1) *c1 = L.c1;
    if (c2) *c2 = L.c2;
    if (cout) *cout = L.cout;
    if (up) *up = L.up;
    if (offset) *offset = L.w_off;
    return FORNOS_OK;
}

fornos_status fornos_denoiser_get_params(const fornos_denoiser* d, float* host, uint64_t count)
{
    if (!d || !host || count != d->n_params)
        return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_get_params: count must equal the parameter count");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host, d->params, sizeof(float) * count, hipMemcpyDeviceToHost));
    return FORNOS_OK;
}

fornos_status fornos_denoiser_set_params(fornos_denoiser* d, const float* host, uint64_t count)
{
    if (!d || !host || count != d->n_params)
        return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_set_params: count must equal the parameter count");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(d->params, host, sizeof(float) * count, hipMemcpyHostToDevice));
    d->refresh_weights();
    d->step = 0;
    launch_1d(k_zero, d->n_params, d->st, d->m, (size_t)d->n_params);
    launch_1d(k_zero, d->n_params, d->st, d->v, (size_t)d->n_params);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return FORNOS_OK;
}

fornos_status fornos_denoiser_forward(fornos_denoiser* d, const float* input, uint32_t n, uint32_t h, uint32_t w,
                                      float* output, float* ms)
{
    if (!d || !input || !output) return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_forward: null argument");
    if (fornos_status s = check_image(d, n, h, w)) return s;
    if (fornos_status s = d->ensure(n, h, w)) return s;
    Timer tm(d->st);
    if (fornos_status s = tm.start()) return s;
    d->forward(input, n, h, w, output, false);
    HIP_TRY(hipGetLastError());
    if (fornos_status s = tm.stop(ms)) return s;
    HIP_TRY(hipDeviceSynchronize());
    return FORNOS_OK;
}

fornos_status fornos_denoiser_backward(fornos_denoiser* d, const float* input, const float* target, uint32_t n,
                                       uint32_t h, uint32_t w, float* output, float* loss)
{
    if (!d || !input || !target) return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_backward: null argument");
    if (fornos_status s = check_image(d, n, h, w)) return s;
    if (fornos_status s = d->ensure(n, h, w)) return s;
    const uint64_t P = (uint64_t)n * h * w;
    hipStream_t st = d->st;
    d->forward(input, n, h, w, d->out, true);
    launch_1d(k_zero_u32, 1, st, d->valid, (size_t)1);
    hipLaunchKernelGGL(k_count_valid, dim3(std::min<uint32_t>(1024, div_up(P, kThreads))), dim3(kThreads), 0, st,
                       input, (uint32_t)FORNOS_DN_IN, (uint32_t)FORNOS_DN_MASK_CHANNEL, (size_t)P, d->valid);
    uint32_t nvalid = 0;
    HIP_TRY(hipMemcpy(&nvalid, d->valid, sizeof(uint32_t), hipMemcpyDeviceToHost));
    const float inv_valid = nvalid ? 1.0f / (float)nvalid : 0.0f;
    launch_1d(k_zero, d->n_params, st, d->grads, (size_t)d->n_params);
    launch_1d(k_zero, 1, st, d->loss, (size_t)1);
    hipLaunchKernelGGL(k_head_back, dim3(std::min<uint32_t>(512, div_up(P, kThreads))), dim3(kThreads), 0, st,
                       (const uint16_t*)d->act[d->dec(0, 1)], input, (uint32_t)FORNOS_DN_IN,
                       (uint32_t)FORNOS_DN_MASK_CHANNEL, (const float*)d->out, target,
                       (const float*)(d->params + d->head_w), inv_valid, d->gact[d->dec(0, 1)],
                       d->grads + d->head_w, d->grads + d->head_b, d->loss, (size_t)P);
    for (uint32_t l = 0; l + 1 < d->levels; ++l) {
        if (fornos_status e = d->conv_backward(d->dec(l, 1), n, h, w)) return e;
        if (fornos_status e = d->conv_backward(d->dec(l, 0), n, h, w)) return e;
    }
    for (int l = (int)d->levels - 1; l >= 0; --l) {
        if (fornos_status e = d->conv_backward(d->enc(l, 1), n, h, w)) return e;
        if (fornos_status e = d->conv_backward(d->enc(l, 0), n, h, w)) return e;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    if (loss) HIP_TRY(hipMemcpy(loss, d->loss, sizeof(float), hipMemcpyDeviceToHost));
    if (output) HIP_TRY(hipMemcpy(output, d->out, sizeof(float) * P * kOut, hipMemcpyDeviceToDevice));
    return FORNOS_OK;
}

fornos_status fornos_denoiser_get_grads(const fornos_denoiser* d, float* host, uint64_t count)
{
    if (!d || !host || count != d->n_params)
        return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_get_grads: count must equal the parameter count");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host, d->grads, sizeof(float) * count, hipMemcpyDeviceToHost));
    return FORNOS_OK;
}

fornos_status fornos_denoiser_adam(fornos_denoiser* d, float lr, float beta1, float beta2, float eps)
{
    if (!d || !(lr > 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return set_error(FORNOS_INVALID_ARG, "fornos_denoiser_adam: bad hyper-parameters");
    d->step += 1;
    const float c1 = 1.0f - std::pow(beta1, (float)d->step), c2 = 1.0f - std::pow(beta2, (float)d->step);
    launch_1d(k_adam, d->n_params, d->st, d->params, (const float*)d->grads, d->m, d->v, d->wb, (size_t)d->n_params,
              lr, beta1, beta2, eps, c1, c2);
    for (const Layer& L : d->layers)
        launch_1d(k_flip_weights, (size_t)L.cout * 9 * (L.c1 + L.c2), d->st, (const uint16_t*)(d->wb + L.w_off),
                  d->wflip + L.f_off, L.cout, L.c1 + L.c2);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return FORNOS_OK;
}

void fornos_denoiser_destroy(fornos_denoiser* d) { delete d; }

fornos_status fornos_dn_conv3x3(const uint16_t* a, uint32_t c1, int up_a, const uint16_t* b, uint32_t c2, uint32_t n,
                                uint32_t h, uint32_t w, const uint16_t* weights, const float* bias, uint32_t cout,
                                int relu, float* out)
{
    if (!a || !weights || !out || (c2 && !b)) return set_error(FORNOS_INVALID_ARG, "fornos_dn_conv3x3: null argument");
    if (!conv_shape_ok(c1, c2, cout) || n == 0 || h == 0 || w == 0 || (up_a && (h % 2 || w % 2)))
        return set_error(FORNOS_INVALID_ARG, "fornos_dn_conv3x3: channels must be multiples of 32, sides even when upsampling");
    if ((uint64_t)n * h * w >= (1ull << 31)) return set_error(FORNOS_UNSUPPORTED, "fornos_dn_conv3x3: too many pixels");
    ConvArgs A{};
    A.a = a;
    A.b = b;
    A.wts = weights;
    A.bias = bias;
    A.out = out;
    A.c1 = c1;
    A.c2 = c2;
    A.cout = cout;
    A.split = cout;
    A.n = n;
    A.h = h;
    A.w = w;
    A.up_a = up_a ? 1 : 0;
    A.relu = relu ? 1 : 0;
    A.out_f32 = 1;
    launch_conv(A, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return FORNOS_OK;
}

fornos_status fornos_dn_conv3x3_wgrad(const uint16_t* a, uint32_t c1, int up_a, const uint16_t* b, uint32_t c2,
                                      uint32_t n, uint32_t h, uint32_t w, const uint16_t* g, uint32_t cout, float* dw,
                                      float* db)
{
    if (!a || !g || !dw || (c2 && !b)) return set_error(FORNOS_INVALID_ARG, "fornos_dn_conv3x3_wgrad: null argument");
    if (!conv_shape_ok(c1, c2, cout) || n == 0 || h == 0 || w == 0 || (up_a && (h % 2 || w % 2)))
        return set_error(FORNOS_INVALID_ARG, "fornos_dn_conv3x3_wgrad: channels must be multiples of 32, sides even when upsampling");
    if ((uint64_t)n * h * w >= (1ull << 31)) return set_error(FORNOS_UNSUPPORTED, "fornos_dn_conv3x3_wgrad: too many pixels");
    const size_t nw = (size_t)cout * 9 * (c1 + c2);
    launch_1d(k_zero, nw, nullptr, dw, nw);
    if (db) launch_1d(k_zero, (size_t)cout, nullptr, db, (size_t)cout);
    WgradArgs W{};
    W.a = a;
    W.b = b;
    W.g = g;
    W.dw = dw;
    W.db = db;
    W.c1 = c1;
    W.c2 = c2;
    W.cout = cout;
    W.n = n;
    W.h = h;
    W.w = w;
    W.up_a = up_a ? 1 : 0;
    float* part = nullptr;
    size_t cap = 0;
    fornos_status e = launch_wgrad(W, nullptr, part, cap);
    if (!e && hipDeviceSynchronize() != hipSuccess) e = set_error(FORNOS_HIP_ERROR, "fornos_dn_conv3x3_wgrad: device error");
    dev_free(part);
    return e;
}

}  // extern "C"
