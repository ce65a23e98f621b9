/*
 * denoise_oracle.c -- CPU oracle of the U-Net denoiser: TEST INFRASTRUCTURE.
 *
 * The denoiser has no counterpart in the reference (it bakes AO and bent
 * normals by brute force only, Src/solver_ao.cpp, Src/solver_bentnormals.cpp),
 * so this file restates the network the library implements, in double
 * precision, from its definition:
 *
 *   encoder level l:  E[l][0] = relu(conv3x3(l ? avgpool2(E[l-1][1]) : x))
 *                     E[l][1] = relu(conv3x3(E[l][0]))
 *   decoder level l = L-2 .. 0 (below = E[L-1][1] or D[l+1][1]):
 *                     D[l][0] = relu(conv3x3(concat(upsample2(below), E[l][1])))
 *                     D[l][1] = relu(conv3x3(D[l][0]))
 *   head:             out[o] = x[o] + hb[o] + sum_c hw[o][c] D[0][1][c]   (o < 4)
 *   loss:             (1/#mask) sum_{mask} sum_o |out[o] - target[o]|
 *
 * with its gradient by reverse-mode differentiation.  Parameters use the
 * library's layout (per layer: weights [cout][3][3][cin], bias [cout]; then
 * the head's [4][base] weights and 4 biases).  With round = 1 the oracle
 * rounds to bfloat16 (nearest even) exactly where the library stores bf16:
 * the packed input, conv weights, every activation and activation gradient;
 * everything else stays double.  With round = 0 it is the plain function,
 * which finite differences pin.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define DN_IN 12
#define DN_OUT 4
#define DN_MASK 11

static double bf(double x)
{
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double R(int round, double x) { return round ? bf(x) : x; }

/* 3x3 convolution, zero padding; a: c1 channels at (h,w) or (h/2,w/2) when
 * up; b: c2 channels at (h,w); w: [cout][9][c1+c2]. */
void dno_conv3x3(const double* a, uint32_t c1, int up, const double* b, uint32_t c2, uint32_t n, uint32_t h,
                 uint32_t w, const double* wts, const double* bias, uint32_t cout, int relu, double* out)
{
    const uint32_t cin = c1 + c2, hh = h / 2, wh = w / 2;
    for (uint32_t bn = 0; bn < n; ++bn)
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x)
                for (uint32_t co = 0; co < cout; ++co) {
                    double s = bias ? bias[co] : 0.0;
                    for (int t = 0; t < 9; ++t) {
                        const int sy = (int)y + t / 3 - 1, sx = (int)x + t % 3 - 1;
                        if (sy < 0 || sx < 0 || sy >= (int)h || sx >= (int)w) continue;
                        const double* wr = wts + ((size_t)co * 9 + t) * cin;
                        const double* pa = up ? a + (((size_t)bn * hh + sy / 2) * wh + sx / 2) * c1
                                              : a + (((size_t)bn * h + sy) * w + sx) * c1;
                        for (uint32_t c = 0; c < c1; ++c) s += wr[c] * pa[c];
                        if (c2) {
                            const double* pb = b + (((size_t)bn * h + sy) * w + sx) * c2;
                            for (uint32_t c = 0; c < c2; ++c) s += wr[c1 + c] * pb[c];
                        }
                    }
                    if (relu && s < 0.0) s = 0.0;
                    out[(((size_t)bn * h + y) * w + x) * cout + co] = s;
                }
}

/* dw[co][t][ci] = sum_p g[p][co] x_t[p][ci]; db[co] = sum_p g[p][co] (overwritten) */
void dno_conv3x3_wgrad(const double* a, uint32_t c1, int up, const double* b, uint32_t c2, uint32_t n, uint32_t h,
                       uint32_t w, const double* g, uint32_t cout, double* dw, double* db)
{
    const uint32_t cin = c1 + c2, hh = h / 2, wh = w / 2;
    memset(dw, 0, sizeof(double) * cout * 9 * cin);
    if (db) memset(db, 0, sizeof(double) * cout);
    for (uint32_t bn = 0; bn < n; ++bn)
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) {
                const double* gp = g + (((size_t)bn * h + y) * w + x) * cout;
                for (uint32_t co = 0; co < cout; ++co) {
                    const double gv = gp[co];
                    if (db) db[co] += gv;
                    if (gv == 0.0) continue;
                    for (int t = 0; t < 9; ++t) {
                        const int sy = (int)y + t / 3 - 1, sx = (int)x + t % 3 - 1;
                        if (sy < 0 || sx < 0 || sy >= (int)h || sx >= (int)w) continue;
                        double* dr = dw + ((size_t)co * 9 + t) * cin;
                        const double* pa = up ? a + (((size_t)bn * hh + sy / 2) * wh + sx / 2) * c1
                                              : a + (((size_t)bn * h + sy) * w + sx) * c1;
                        for (uint32_t c = 0; c < c1; ++c) dr[c] += gv * pa[c];
                        if (c2) {
                            const double* pb = b + (((size_t)bn * h + sy) * w + sx) * c2;
                            for (uint32_t c = 0; c < c2; ++c) dr[c1 + c] += gv * pb[c];
                        }
                    }
                }
            }
}

/* gradient of the conv input(s): ga [..][c1] at full resolution (the
 * upsampled tensor's gradient before the 2x2 sum), gb [..][c2] */
static void conv_input_grad(const double* g, uint32_t cout, uint32_t n, uint32_t h, uint32_t w, const double* wts,
                            uint32_t c1, uint32_t c2, double* ga, double* gb)
{
    const uint32_t cin = c1 + c2;
    memset(ga, 0, sizeof(double) * n * h * w * c1);
    if (c2) memset(gb, 0, sizeof(double) * n * h * w * c2);
    for (uint32_t bn = 0; bn < n; ++bn)
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) {
                const double* gp = g + (((size_t)bn * h + y) * w + x) * cout;
                for (int t = 0; t < 9; ++t) {
                    const int sy = (int)y + t / 3 - 1, sx = (int)x + t % 3 - 1;
                    if (sy < 0 || sx < 0 || sy >= (int)h || sx >= (int)w) continue;
                    const size_t sp = ((size_t)bn * h + sy) * w + sx;
                    for (uint32_t co = 0; co < cout; ++co) {
                        const double gv = gp[co];
                        if (gv == 0.0) continue;
                        const double* wr = wts + ((size_t)co * 9 + t) * cin;
                        for (uint32_t c = 0; c < c1; ++c) ga[sp * c1 + c] += gv * wr[c];
                        for (uint32_t c = 0; c < c2; ++c) gb[sp * c2 + c] += gv * wr[c1 + c];
                    }
                }
            }
}

typedef struct {
    uint32_t c1, c2, cout, up, level;
    size_t w_off, b_off;
} Layer;

typedef struct {
    uint32_t levels, base, nl;
    Layer L[16];
    size_t head_w, head_b, n_params;
} Net;

static void net_init(Net* N, uint32_t levels, uint32_t base)
{
    size_t off = 0;
    N->levels = levels;
    N->base = base;
    N->nl = 0;
#define ADD(C1, C2, CO, UP, LV)                                                                          \
    do {                                                                                                 \
        Layer* l_ = &N->L[N->nl++];                                                                      \
        l_->c1 = (C1); l_->c2 = (C2); l_->cout = (CO); l_->up = (UP); l_->level = (LV); l_->w_off = off; \
        off += (size_t)(CO) * 9 * ((C1) + (C2));                                                         \
        l_->b_off = off;                                                                                 \
        off += (CO);                                                                                     \
    } while (0)
    for (uint32_t l = 0; l < levels; ++l) {
        ADD(l == 0 ? 32u : base << (l - 1), 0u, base << l, 0u, l);
        ADD(base << l, 0u, base << l, 0u, l);
    }
    for (int l = (int)levels - 2; l >= 0; --l) {
        ADD(base << (l + 1), base << l, base << l, 1u, (uint32_t)l);
        ADD(base << l, 0u, base << l, 0u, (uint32_t)l);
    }
#undef ADD
    N->head_w = off;
    off += (size_t)DN_OUT * base;
    N->head_b = off;
    off += DN_OUT;
    N->n_params = off;
}

uint64_t dno_param_count(uint32_t levels, uint32_t base)
{
    Net N;
    net_init(&N, levels, base);
    return N.n_params;
}

static uint32_t enc(const Net* N, uint32_t l, uint32_t k) { return 2 * l + k; }
static uint32_t dec(const Net* N, uint32_t l, uint32_t k) { return 2 * N->levels + 2 * (N->levels - 2 - l) + k; }

typedef struct {
    double* x0;
    double* act[16];
    double* pool[8];
    double* wr[16]; /* (rounded) weights per layer */
} Fwd;

static size_t px_at(uint32_t n, uint32_t h, uint32_t w, uint32_t l) { return (size_t)n * (h >> l) * (w >> l); }

static void layer_inputs(const Net* N, const Fwd* F, uint32_t i, const double** a, const double** b)
{
    *b = NULL;
    if (i < 2 * N->levels) {
        const uint32_t l = i / 2;
        *a = (i & 1) ? F->act[i - 1] : (l == 0 ? F->x0 : F->pool[l]);
    } else {
        const uint32_t j = i - 2 * N->levels, l = N->levels - 2 - j / 2;
        if (j & 1) {
            *a = F->act[i - 1];
        } else {
            *a = (l == N->levels - 2) ? F->act[enc(N, N->levels - 1, 1)] : F->act[dec(N, l + 1, 1)];
            *b = F->act[enc(N, l, 1)];
        }
    }
}

static void conv_fwd(const Net* N, Fwd* F, const double* params, uint32_t i, uint32_t n, uint32_t h, uint32_t w,
                     int round)
{
    const Layer* L = &N->L[i];
    const double *a, *b;
    layer_inputs(N, F, i, &a, &b);
    const uint32_t hl = h >> L->level, wl = w >> L->level;
    const size_t P = px_at(n, h, w, L->level);
    dno_conv3x3(a, L->c1, (int)L->up, b, L->c2, n, hl, wl, F->wr[i], params + L->b_off, L->cout, 1, F->act[i]);
    for (size_t k = 0; k < P * L->cout; ++k) F->act[i][k] = R(round, F->act[i][k]);
}

static void forward(const Net* N, Fwd* F, const double* params, const double* in, uint32_t n, uint32_t h, uint32_t w,
                    int round, double* out)
{
    const size_t P = (size_t)n * h * w;
    for (size_t p = 0; p < P; ++p)
        for (uint32_t c = 0; c < 32; ++c) F->x0[p * 32 + c] = c < DN_IN ? R(round, in[p * DN_IN + c]) : 0.0;
    for (uint32_t i = 0; i < N->nl; ++i) {
        const Layer* L = &N->L[i];
        const size_t nw = (size_t)L->cout * 9 * (L->c1 + L->c2);
        for (size_t k = 0; k < nw; ++k) F->wr[i][k] = R(round, params[L->w_off + k]);
    }
    for (uint32_t l = 0; l < N->levels; ++l) {
        if (l > 0) {
            const uint32_t c = N->base << (l - 1), hi = h >> (l - 1), wi = w >> (l - 1);
            const double* src = F->act[enc(N, l - 1, 1)];
            for (uint32_t bn = 0; bn < n; ++bn)
                for (uint32_t y = 0; y < hi / 2; ++y)
                    for (uint32_t x = 0; x < wi / 2; ++x)
                        for (uint32_t k = 0; k < c; ++k) {
                            double s = 0.0;
                            for (int q = 0; q < 4; ++q)
                                s += src[(((size_t)bn * hi + 2 * y + q / 2) * wi + 2 * x + q % 2) * c + k];
                            F->pool[l][(((size_t)bn * (hi / 2) + y) * (wi / 2) + x) * c + k] = R(round, 0.25 * s);
                        }
        }
        conv_fwd(N, F, params, enc(N, l, 0), n, h, w, round);
        conv_fwd(N, F, params, enc(N, l, 1), n, h, w, round);
    }
    for (int l = (int)N->levels - 2; l >= 0; --l) {
        conv_fwd(N, F, params, dec(N, (uint32_t)l, 0), n, h, w, round);
        conv_fwd(N, F, params, dec(N, (uint32_t)l, 1), n, h, w, round);
    }
    const double* y = F->act[dec(N, 0, 1)];
    const double *hw = params + N->head_w, *hb = params + N->head_b;
    for (size_t p = 0; p < P; ++p)
        for (uint32_t o = 0; o < DN_OUT; ++o) {
            double s = in[p * DN_IN + o] + hb[o];
            for (uint32_t c = 0; c < N->base; ++c) s += hw[o * N->base + c] * y[p * N->base + c];
            out[p * DN_OUT + o] = s;
        }
}

static int fwd_alloc(const Net* N, Fwd* F, uint32_t n, uint32_t h, uint32_t w)
{
    memset(F, 0, sizeof(*F));
    F->x0 = (double*)malloc(sizeof(double) * (size_t)n * h * w * 32);
    for (uint32_t i = 0; i < N->nl; ++i) {
        const Layer* L = &N->L[i];
        F->act[i] = (double*)malloc(sizeof(double) * px_at(n, h, w, L->level) * L->cout);
        F->wr[i] = (double*)malloc(sizeof(double) * L->cout * 9 * (L->c1 + L->c2));
    }
    for (uint32_t l = 1; l < N->levels; ++l)
        F->pool[l] = (double*)malloc(sizeof(double) * px_at(n, h, w, l) * (N->base << (l - 1)));
    return F->x0 != NULL;
}

static void fwd_free(const Net* N, Fwd* F)
{
    free(F->x0);
    for (uint32_t i = 0; i < N->nl; ++i) {
        free(F->act[i]);
        free(F->wr[i]);
    }
    for (uint32_t l = 1; l < N->levels; ++l) free(F->pool[l]);
}

int dno_forward(uint32_t levels, uint32_t base, const double* params, const double* in, uint32_t n, uint32_t h,
                uint32_t w, int round, double* out)
{
    Net N;
    Fwd F;
    net_init(&N, levels, base);
    if (N.nl > 16 || !fwd_alloc(&N, &F, n, h, w)) return -1;
    forward(&N, &F, params, in, n, h, w, round, out);
    fwd_free(&N, &F);
    return 0;
}

/* loss and its gradient with respect to every parameter */
int dno_backward(uint32_t levels, uint32_t base, const double* params, const double* in, const double* target,
                 uint32_t n, uint32_t h, uint32_t w, int round, double* grads, double* loss, double* out)
{
    Net N;
    Fwd F;
    net_init(&N, levels, base);
    if (N.nl > 16 || !fwd_alloc(&N, &F, n, h, w)) return -1;
    const size_t P = (size_t)n * h * w;
    double* o = (double*)malloc(sizeof(double) * P * DN_OUT);
    forward(&N, &F, params, in, n, h, w, round, o);
    if (out) memcpy(out, o, sizeof(double) * P * DN_OUT);
    memset(grads, 0, sizeof(double) * N.n_params);

    double* gact[16];
    for (uint32_t i = 0; i < N.nl; ++i) gact[i] = (double*)calloc(px_at(n, h, w, N.L[i].level) * N.L[i].cout, sizeof(double));

    /* loss and head */
    size_t nvalid = 0;
    for (size_t p = 0; p < P; ++p) nvalid += in[p * DN_IN + DN_MASK] > 0.5;
    const double inv = nvalid ? 1.0 / (double)nvalid : 0.0;
    double lsum = 0.0;
    const double* y = F.act[dec(&N, 0, 1)];
    const double* hw = params + N.head_w;
    double* gy = gact[dec(&N, 0, 1)];
    for (size_t p = 0; p < P; ++p) {
        double d[DN_OUT] = {0, 0, 0, 0};
        if (in[p * DN_IN + DN_MASK] > 0.5)
            for (uint32_t k = 0; k < DN_OUT; ++k) {
                const double e = o[p * DN_OUT + k] - target[p * DN_OUT + k];
                lsum += fabs(e);
                d[k] = e > 0 ? inv : (e < 0 ? -inv : 0.0);
            }
        for (uint32_t k = 0; k < DN_OUT; ++k) {
            grads[N.head_b + k] += d[k];
            for (uint32_t c = 0; c < N.base; ++c) grads[N.head_w + k * N.base + c] += d[k] * y[p * N.base + c];
        }
        for (uint32_t c = 0; c < N.base; ++c) {
            double s = 0.0;
            for (uint32_t k = 0; k < DN_OUT; ++k) s += d[k] * hw[k * N.base + c];
            gy[p * N.base + c] = y[p * N.base + c] > 0 ? R(round, s) : 0.0;
        }
    }
    *loss = lsum * inv;

    /* layers, in the library's order */
    uint32_t order[16], no = 0;
    for (uint32_t l = 0; l + 1 < N.levels; ++l) {
        order[no++] = dec(&N, l, 1);
        order[no++] = dec(&N, l, 0);
    }
    for (int l = (int)N.levels - 1; l >= 0; --l) {
        order[no++] = enc(&N, (uint32_t)l, 1);
        order[no++] = enc(&N, (uint32_t)l, 0);
    }
    for (uint32_t k = 0; k < no; ++k) {
        const uint32_t i = order[k];
        const Layer* L = &N.L[i];
        const uint32_t hl = h >> L->level, wl = w >> L->level;
        const size_t Pl = px_at(n, h, w, L->level);
        const double *a, *b;
        layer_inputs(&N, &F, i, &a, &b);
        const uint32_t cin = L->c1 + L->c2;
        double* dw = (double*)malloc(sizeof(double) * L->cout * 9 * cin);
        dno_conv3x3_wgrad(a, L->c1, (int)L->up, b, L->c2, n, hl, wl, gact[i], L->cout, dw, grads + L->b_off);
        memcpy(grads + L->w_off, dw, sizeof(double) * L->cout * 9 * cin);
        free(dw);
        if (i == 0) continue;
        double* ga = (double*)malloc(sizeof(double) * Pl * L->c1);
        double* gb = L->c2 ? (double*)malloc(sizeof(double) * Pl * L->c2) : NULL;
        conv_input_grad(gact[i], L->cout, n, hl, wl, F.wr[i], L->c1, L->c2, ga, gb);
        if (i < 2 * N.levels) {
            const uint32_t l = i / 2;
            if (i & 1) {
                const double* m = F.act[i - 1];
                for (size_t q = 0; q < Pl * L->c1; ++q) gact[i - 1][q] = m[q] > 0 ? R(round, ga[q]) : 0.0;
            } else { /* through the pool into E[l-1][1] (accumulate) */
                const uint32_t c = L->c1, hi = hl * 2, wi = wl * 2;
                const uint32_t src = enc(&N, l - 1, 1);
                for (size_t q = 0; q < Pl * c; ++q) ga[q] = R(round, ga[q]);
                for (uint32_t bn = 0; bn < n; ++bn)
                    for (uint32_t yy = 0; yy < hi; ++yy)
                        for (uint32_t xx = 0; xx < wi; ++xx)
                            for (uint32_t ch = 0; ch < c; ++ch) {
                                const size_t di = (((size_t)bn * hi + yy) * wi + xx) * c + ch;
                                const size_t si = (((size_t)bn * hl + yy / 2) * wl + xx / 2) * c + ch;
                                if (F.act[src][di] > 0) gact[src][di] = R(round, gact[src][di] + 0.25 * ga[si]);
                            }
            }
        } else {
            const uint32_t j = i - 2 * N.levels, l = N.levels - 2 - j / 2;
            if (j & 1) {
                const double* m = F.act[i - 1];
                for (size_t q = 0; q < Pl * L->c1; ++q) gact[i - 1][q] = m[q] > 0 ? R(round, ga[q]) : 0.0;
            } else {
                const uint32_t sk = enc(&N, l, 1);
                for (size_t q = 0; q < Pl * L->c2; ++q) gact[sk][q] = F.act[sk][q] > 0 ? R(round, gb[q]) : 0.0;
                const uint32_t bl = (l == N.levels - 2) ? enc(&N, N.levels - 1, 1) : dec(&N, l + 1, 1);
                const uint32_t c = L->c1, hs = hl / 2, ws = wl / 2;
                for (size_t q = 0; q < Pl * c; ++q) ga[q] = R(round, ga[q]);
                for (uint32_t bn = 0; bn < n; ++bn)
                    for (uint32_t yy = 0; yy < hs; ++yy)
                        for (uint32_t xx = 0; xx < ws; ++xx)
                            for (uint32_t ch = 0; ch < c; ++ch) {
                                double s = 0.0;
                                for (int q = 0; q < 4; ++q)
                                    s += ga[(((size_t)bn * hl + 2 * yy + q / 2) * wl + 2 * xx + q % 2) * c + ch];
                                const size_t di = (((size_t)bn * hs + yy) * ws + xx) * c + ch;
                                gact[bl][di] = F.act[bl][di] > 0 ? R(round, s) : 0.0;
                            }
            }
        }
        free(ga);
        free(gb);
    }
    for (uint32_t i = 0; i < N.nl; ++i) free(gact[i]);
    free(o);
    fwd_free(&N, &F);
    return 0;
}
