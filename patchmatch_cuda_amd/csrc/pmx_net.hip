/*
 * pmx_net.hip -- the proposal network: a 4-level BF16 encoder-decoder of 3x3
 * convolutions (channels 32/64/128/256) on the MFMA kernels of pmx_cnn.hip,
 * with its training step (L1 on inverse depth + L1 of the predicted error
 * scale, Adam) and the C ABI declared in include/pmx.h.
 *
 * Input (1/4 resolution, NHWC, 64 channels): 32 plane-sweep costs + the
 * reference intensity + 31 zero channels.  Output (head, FP32, 32 channels):
 * channel 0 = inverse-depth logit, 1 = confidence logit, the rest unused.
 *
 *   e0a 64->32  e0b 32->32              level 0 (H0 x W0)
 *   e1a 32->64 /2  e1b 64->64           level 1
 *   e2a 64->128 /2 e2b 128->128         level 2
 *   e3a 128->256 /2 e3b 256->256        level 3
 *   d2a [up(e3b), e2b] 384->128, d2b    level 2
 *   d1a [up(d2b), e1b] 192->64,  d1b    level 1
 *   d0a [up(d1b), e0b] 96->32,   d0b    level 0
 *   head 32->32 (no activation)
 * Every convolution but the head is followed by ReLU; "up" is nearest x2.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "pmx_internal.h"

namespace {

constexpr int kInC = 64;   /* input channels (33 used) */
constexpr int kHeadC = 32; /* head channels (2 used) */
constexpr int kLayers = 15;

/* tensor ids: 0..13 = outputs of layers 0..13, then the input and the three
 * concatenations */
enum { T_X = 14, T_C2, T_C1, T_C0, T_COUNT };

struct LayerDef {
    int cin, cout, stride, relu;
    int in, level_in;      /* input tensor and its level */
    int gin, accumulate;   /* gradient of the input (-1: none) */
};

const LayerDef kDefs[kLayers] = {
    {64, 32, 1, 1, T_X, 0, -1, 0},     /* e0a */
    {32, 32, 1, 1, 0, 0, 0, 0},        /* e0b */
    {32, 64, 2, 1, 1, 0, 1, 1},        /* e1a: e0b's gradient also has the skip part */
    {64, 64, 1, 1, 2, 1, 2, 0},        /* e1b */
    {64, 128, 2, 1, 3, 1, 3, 1},       /* e2a */
    {128, 128, 1, 1, 4, 2, 4, 0},      /* e2b */
    {128, 256, 2, 1, 5, 2, 5, 1},      /* e3a */
    {256, 256, 1, 1, 6, 3, 6, 0},      /* e3b */
    {384, 128, 1, 1, T_C2, 2, T_C2, 0},/* d2a */
    {128, 128, 1, 1, 8, 2, 8, 0},      /* d2b */
    {192, 64, 1, 1, T_C1, 1, T_C1, 0}, /* d1a */
    {64, 64, 1, 1, 10, 1, 10, 0},      /* d1b */
    {96, 32, 1, 1, T_C0, 0, T_C0, 0},  /* d0a */
    {32, 32, 1, 1, 12, 0, 12, 0},      /* d0b */
    {32, kHeadC, 1, 0, 13, 0, 13, 0},  /* head */
};

/* the concatenations: (output tensor, upsampled tensor, skip tensor, level of the output) */
struct CatDef {
    int out, up, skip, level;
};
const CatDef kCats[3] = {{T_C2, 7, 5, 2}, {T_C1, 9, 3, 1}, {T_C0, 11, 1, 0}};

int out_level(int i) { return kDefs[i].level_in + (kDefs[i].stride == 2 ? 1 : 0); }

int tensor_level(int t) {
    if (t < kLayers - 1) return out_level(t);
    if (t == T_X) return 0;
    for (const CatDef& c : kCats)
        if (c.out == t) return c.level;
    return 0;
}

int tensor_channels(int t) {
    if (t < kLayers - 1) return kDefs[t].cout;
    if (t == T_X) return kInC;
    for (const CatDef& c : kCats)
        if (c.out == t) return kDefs[c.up].cout + kDefs[c.skip].cout;
    return 0;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double gauss(uint64_t& s) {
    const double u1 = ((splitmix(s) >> 11) + 1) * (1.0 / 9007199254740993.0);
    const double u2 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

}  // namespace

struct pmx_net {
    pmx_ctx* ctx = nullptr;
    int max_batch = 0, H0 = 0, W0 = 0;
    float inv_lo = 0.0f, inv_hi = 1.0f; /* inverse-depth range of the head */
    size_t w_off[kLayers] = {}, b_off[kLayers] = {};
    size_t n_w = 0, n_total = 0;
    float* master = nullptr;   /* FP32 weights then biases */
    uint16_t* wb = nullptr;    /* bf16 weights [Cout][9][Cin] */
    uint16_t* wt = nullptr;    /* bf16 weights [Cin][9][Cout] (backward data) */
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int adam_t = 0;
    uint16_t* act[T_COUNT] = {};
    uint16_t* gact[T_COUNT] = {};
    float* head = nullptr;     /* FP32 head output */
    uint16_t* ghead = nullptr; /* its gradient (bf16) */
    float* sums = nullptr;     /* loss sums (2) */
    size_t cap[T_COUNT] = {};  /* elements */
    /* shape of the last forward */
    int N = 0, fH0 = 0, fW0 = 0;
    bool wt_fresh = false;
};

namespace {

void level_dims(int H0, int W0, int l, int& H, int& W) {
    H = H0;
    W = W0;
    for (int i = 0; i < l; i++) {
        H = (H - 1) / 2 + 1;
        W = (W - 1) / 2 + 1;
    }
}

#define NCHK(net, expr)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return pmx_fail((net)->ctx, PMX_ERR_HIP, "%s in %s at line %d",                  \
                            hipGetErrorString(e_), __FILE__, __LINE__);                      \
    } while (0)

int net_alloc(pmx_net* net, void** p, size_t bytes) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return pmx_fail(net->ctx, PMX_ERR_ALLOC, "device allocation of %zu bytes failed", bytes);
    }
    return PMX_OK;
}

void net_free_all(pmx_net* net) {
    void* ps[] = {net->master, net->wb, net->wt, net->grad, net->adam_m, net->adam_v,
                  net->head, net->ghead, net->sums};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (int t = 0; t < T_COUNT; t++) {
        if (net->act[t]) (void)hipFree(net->act[t]);
        if (net->gact[t]) (void)hipFree(net->gact[t]);
    }
}

int check_shape(pmx_net* net, int N, int H0, int W0) {
    if (N <= 0 || H0 <= 0 || W0 <= 0)
        return pmx_fail(net->ctx, PMX_ERR_ARG, "batch and sizes must be positive");
    for (int t = 0; t < T_COUNT; t++) {
        int H, W;
        level_dims(H0, W0, tensor_level(t), H, W);
        if ((size_t)N * H * W * tensor_channels(t) > net->cap[t])
            return pmx_fail(net->ctx, PMX_ERR_ARG,
                            "batch %d of %dx%d exceeds the network's capacity (%d of %dx%d)", N,
                            W0, H0, net->max_batch, net->W0, net->H0);
    }
    return PMX_OK;
}

int refresh_wt(pmx_net* net) {
    hipStream_t s = net->ctx->stream;
    for (int i = 0; i < kLayers; i++)
        NCHK(net, pmx_launch_weight_t(s, net->wb + net->w_off[i], net->wt + net->w_off[i],
                                      kDefs[i].cout, kDefs[i].cin));
    net->wt_fresh = true;
    return PMX_OK;
}

int run_forward(pmx_net* net, const uint16_t* x, int N, int H0, int W0) {
    int rc = check_shape(net, N, H0, W0);
    if (rc != PMX_OK) return rc;
    hipStream_t s = net->ctx->stream;
    net->N = N;
    net->fH0 = H0;
    net->fW0 = W0;
    const uint16_t* in_ptr[T_COUNT];
    for (int t = 0; t < T_COUNT; t++) in_ptr[t] = net->act[t];
    in_ptr[T_X] = x;
    auto conv = [&](int i) -> int {
        const LayerDef& L = kDefs[i];
        int H, W;
        level_dims(H0, W0, L.level_in, H, W);
        const bool head = i == kLayers - 1;
        NCHK(net, pmx_launch_conv_fwd(s, in_ptr[L.in], net->wb + net->w_off[i],
                                      net->master + net->b_off[i], head ? nullptr : net->act[i],
                                      head ? net->head : nullptr, N, H, W, L.cin, L.cout, L.stride,
                                      L.relu));
        return PMX_OK;
    };
    auto cat = [&](const CatDef& c) -> int {
        int H, W, Ha, Wa;
        level_dims(H0, W0, c.level, H, W);
        level_dims(H0, W0, c.level + 1, Ha, Wa);
        NCHK(net, pmx_launch_upcat(s, net->act[c.up], net->act[c.skip], net->act[c.out], N, H, W,
                                   Ha, Wa, kDefs[c.up].cout, kDefs[c.skip].cout));
        return PMX_OK;
    };
    for (int i = 0; i <= 7; i++)
        if ((rc = conv(i)) != PMX_OK) return rc;
    if ((rc = cat(kCats[0])) != PMX_OK) return rc;
    for (int i = 8; i <= 9; i++)
        if ((rc = conv(i)) != PMX_OK) return rc;
    if ((rc = cat(kCats[1])) != PMX_OK) return rc;
    for (int i = 10; i <= 11; i++)
        if ((rc = conv(i)) != PMX_OK) return rc;
    if ((rc = cat(kCats[2])) != PMX_OK) return rc;
    for (int i = 12; i <= 14; i++)
        if ((rc = conv(i)) != PMX_OK) return rc;
    return PMX_OK;
}

/* gradients of the last forward's head loss, added to net->grad */
int run_backward(pmx_net* net, const uint16_t* x, const float* tgt, float w_conf) {
    hipStream_t s = net->ctx->stream;
    const int N = net->N, H0 = net->fH0, W0 = net->fW0;
    const int P0 = N * H0 * W0;
    if (!net->wt_fresh) {
        int rc = refresh_wt(net);
        if (rc != PMX_OK) return rc;
    }
    NCHK(net, pmx_launch_zero_f32(s, net->sums, 2));
    NCHK(net, pmx_launch_head_loss(s, net->head, tgt, net->ghead, P0, kHeadC, net->inv_lo,
                                   net->inv_hi, 1.0f / (float)P0, w_conf, net->sums));
    const uint16_t* in_ptr[T_COUNT];
    for (int t = 0; t < T_COUNT; t++) in_ptr[t] = net->act[t];
    in_ptr[T_X] = x;
    for (int i = kLayers - 1; i >= 0; i--) {
        const LayerDef& L = kDefs[i];
        int H, W, Ho, Wo;
        level_dims(H0, W0, L.level_in, H, W);
        level_dims(H0, W0, out_level(i), Ho, Wo);
        uint16_t* dZ = (i == kLayers - 1) ? net->ghead : net->gact[i];
        const size_t nz = (size_t)N * Ho * Wo * L.cout;
        if (L.relu) NCHK(net, pmx_launch_relu_bwd(s, dZ, net->act[i], nz));
        NCHK(net, pmx_launch_bias_grad(s, dZ, net->grad + net->b_off[i], N * Ho * Wo, L.cout));
        NCHK(net, pmx_launch_conv_bwd_weight(s, in_ptr[L.in], dZ, net->grad + net->w_off[i], N, H,
                                             W, L.cin, L.cout, L.stride));
        if (L.gin >= 0)
            NCHK(net, pmx_launch_conv_bwd_data(s, dZ, net->wt + net->w_off[i], net->gact[L.gin], N,
                                               H, W, L.cin, L.cout, L.stride, L.accumulate));
        for (const CatDef& c : kCats) {
            if (c.out != L.in) continue;
            int Hc, Wc, Ha, Wa;
            level_dims(H0, W0, c.level, Hc, Wc);
            level_dims(H0, W0, c.level + 1, Ha, Wa);
            NCHK(net, pmx_launch_upcat_bwd(s, net->gact[c.out], net->gact[c.up], net->gact[c.skip],
                                           N, Hc, Wc, Ha, Wa, kDefs[c.up].cout,
                                           kDefs[c.skip].cout));
        }
    }
    return PMX_OK;
}

int copy_sums(pmx_net* net, float* loss) {
    if (!loss) return PMX_OK;
    float h[2];
    NCHK(net, hipMemcpyAsync(h, net->sums, sizeof(h), hipMemcpyDeviceToHost, net->ctx->stream));
    NCHK(net, hipStreamSynchronize(net->ctx->stream));
    const float P0 = (float)net->N * net->fH0 * net->fW0;
    loss[0] = h[0] / P0;
    loss[1] = h[1] / P0;
    return PMX_OK;
}

}  // namespace

pmx_ctx* pmx_net_context(const pmx_net* net) { return net ? net->ctx : nullptr; }

extern "C" {

int pmx_net_create(pmx_ctx* ctx, int max_batch, int H0, int W0, float depth_near, float depth_far,
                   unsigned long long seed, pmx_net** out) {
    if (!ctx || !out) return pmx_fail(ctx, PMX_ERR_ARG, "ctx and out must not be NULL");
    *out = nullptr;
    if (max_batch <= 0 || H0 <= 0 || W0 <= 0)
        return pmx_fail(ctx, PMX_ERR_ARG, "batch and sizes must be positive");
    if (!(depth_near > 0.0f) || !(depth_far > depth_near))
        return pmx_fail(ctx, PMX_ERR_ARG, "need 0 < near < far");
    pmx_net* net = new pmx_net();
    net->ctx = ctx;
    net->max_batch = max_batch;
    net->H0 = H0;
    net->W0 = W0;
    net->inv_lo = 1.0f / depth_far;
    net->inv_hi = 1.0f / depth_near;
    size_t off = 0;
    for (int i = 0; i < kLayers; i++) {
        net->w_off[i] = off;
        off += (size_t)kDefs[i].cout * 9 * kDefs[i].cin;
    }
    net->n_w = off;
    for (int i = 0; i < kLayers; i++) {
        net->b_off[i] = off;
        off += (size_t)kDefs[i].cout;
    }
    net->n_total = off;
    (void)hipSetDevice(ctx->device);
    int rc = PMX_OK;
    auto A = [&](void** p, size_t bytes) {
        if (rc == PMX_OK) rc = net_alloc(net, p, bytes);
    };
    A((void**)&net->master, net->n_total * 4);
    A((void**)&net->wb, net->n_w * 2);
    A((void**)&net->wt, net->n_w * 2);
    A((void**)&net->grad, net->n_total * 4);
    A((void**)&net->adam_m, net->n_total * 4);
    A((void**)&net->adam_v, net->n_total * 4);
    A((void**)&net->sums, 64);
    for (int t = 0; t < T_COUNT; t++) {
        int H, W;
        level_dims(H0, W0, tensor_level(t), H, W);
        net->cap[t] = (size_t)max_batch * H * W * tensor_channels(t);
        if (t != T_X) A((void**)&net->act[t], net->cap[t] * 2);
        A((void**)&net->gact[t], net->cap[t] * 2);
    }
    A((void**)&net->head, (size_t)max_batch * H0 * W0 * kHeadC * 4);
    A((void**)&net->ghead, (size_t)max_batch * H0 * W0 * kHeadC * 2);
    if (rc != PMX_OK) {
        net_free_all(net);
        delete net;
        return rc;
    }
    /* He initialisation from the seed (the first layer's fan-in counts its 33
     * used channels); the head starts small, biases at 0 */
    std::vector<float> p(net->n_total, 0.0f);
    uint64_t st = seed ^ 0x5DEECE66Dull;
    for (int i = 0; i < kLayers; i++) {
        const LayerDef& L = kDefs[i];
        const int fan_in = 9 * (i == 0 ? 33 : L.cin);
        const double sd = (i == kLayers - 1) ? 0.01 : std::sqrt(2.0 / fan_in);
        const size_t n = (size_t)L.cout * 9 * L.cin;
        for (size_t k = 0; k < n; k++) p[net->w_off[i] + k] = (float)(sd * gauss(st));
    }
    rc = pmx_net_set_params(net, p.data(), p.size());
    if (rc == PMX_OK) {
        hipStream_t s = ctx->stream;
        hipError_t e = hipMemsetAsync(net->grad, 0, net->n_total * 4, s);
        if (e == hipSuccess) e = hipMemsetAsync(net->adam_m, 0, net->n_total * 4, s);
        if (e == hipSuccess) e = hipMemsetAsync(net->adam_v, 0, net->n_total * 4, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = pmx_fail(ctx, PMX_ERR_HIP, "%s", hipGetErrorString(e));
    }
    if (rc != PMX_OK) {
        net_free_all(net);
        delete net;
        return rc;
    }
    *out = net;
    return PMX_OK;
}

int pmx_net_destroy(pmx_net* net) {
    if (!net) return PMX_OK;
    (void)hipStreamSynchronize(net->ctx->stream);
    net_free_all(net);
    delete net;
    return PMX_OK;
}

int pmx_net_num_params(const pmx_net* net, size_t* n_total, size_t* n_weights) {
    if (!net) return pmx_fail(nullptr, PMX_ERR_ARG, "net is NULL");
    if (n_total) *n_total = net->n_total;
    if (n_weights) *n_weights = net->n_w;
    return PMX_OK;
}

int pmx_net_layer_offsets(const pmx_net* net, size_t* w_off, size_t* b_off) {
    if (!net || !w_off || !b_off) return pmx_fail(nullptr, PMX_ERR_ARG, "NULL argument");
    for (int i = 0; i < kLayers; i++) {
        w_off[i] = net->w_off[i];
        b_off[i] = net->b_off[i];
    }
    return PMX_OK;
}

int pmx_net_set_params(pmx_net* net, const float* host, size_t n) {
    if (!net || !host) return pmx_fail(net ? net->ctx : nullptr, PMX_ERR_ARG, "NULL argument");
    if (n != net->n_total)
        return pmx_fail(net->ctx, PMX_ERR_ARG, "expected %zu parameters, got %zu", net->n_total, n);
    hipStream_t s = net->ctx->stream;
    NCHK(net, hipMemcpyAsync(net->master, host, n * 4, hipMemcpyHostToDevice, s));
    NCHK(net, pmx_launch_f32_to_bf16(s, net->master, net->wb, net->n_w));
    NCHK(net, hipStreamSynchronize(s));
    net->wt_fresh = false;
    return PMX_OK;
}

int pmx_net_get_params(pmx_net* net, float* host, size_t n) {
    if (!net || !host) return pmx_fail(net ? net->ctx : nullptr, PMX_ERR_ARG, "NULL argument");
    if (n != net->n_total)
        return pmx_fail(net->ctx, PMX_ERR_ARG, "expected %zu parameters, got %zu", net->n_total, n);
    NCHK(net, hipMemcpyAsync(host, net->master, n * 4, hipMemcpyDeviceToHost, net->ctx->stream));
    NCHK(net, hipStreamSynchronize(net->ctx->stream));
    return PMX_OK;
}

int pmx_net_get_grads(pmx_net* net, float* host, size_t n) {
    if (!net || !host) return pmx_fail(net ? net->ctx : nullptr, PMX_ERR_ARG, "NULL argument");
    if (n != net->n_total)
        return pmx_fail(net->ctx, PMX_ERR_ARG, "expected %zu parameters, got %zu", net->n_total, n);
    NCHK(net, hipMemcpyAsync(host, net->grad, n * 4, hipMemcpyDeviceToHost, net->ctx->stream));
    NCHK(net, hipStreamSynchronize(net->ctx->stream));
    return PMX_OK;
}

int pmx_net_zero_grads(pmx_net* net) {
    if (!net) return pmx_fail(nullptr, PMX_ERR_ARG, "net is NULL");
    NCHK(net, pmx_launch_zero_f32(net->ctx->stream, net->grad, net->n_total));
    return PMX_OK;
}

int pmx_net_forward(pmx_net* net, const void* x, int N, int H0, int W0) {
    if (!net || !x) return pmx_fail(net ? net->ctx : nullptr, PMX_ERR_ARG, "NULL argument");
    return run_forward(net, (const uint16_t*)x, N, H0, W0);
}

int pmx_net_get_head(pmx_net* net, float* host, size_t n) {
    if (!net || !host) return pmx_fail(net ? net->ctx : nullptr, PMX_ERR_ARG, "NULL argument");
    const size_t want = (size_t)net->N * net->fH0 * net->fW0 * kHeadC;
    if (net->N == 0 || n != want)
        return pmx_fail(net->ctx, PMX_ERR_ARG, "head holds %zu values, asked for %zu", want, n);
    NCHK(net, hipMemcpyAsync(host, net->head, n * 4, hipMemcpyDeviceToHost, net->ctx->stream));
    NCHK(net, hipStreamSynchronize(net->ctx->stream));
    return PMX_OK;
}

int pmx_net_backward(pmx_net* net, const void* x, const float* tgt, float w_conf, float* loss) {
    if (!net || !x || !tgt) return pmx_fail(net ? net->ctx : nullptr, PMX_ERR_ARG, "NULL argument");
    if (net->N == 0) return pmx_fail(net->ctx, PMX_ERR_STATE, "no forward pass to differentiate");
    int rc = run_backward(net, (const uint16_t*)x, tgt, w_conf);
    if (rc != PMX_OK) return rc;
    return copy_sums(net, loss);
}

int pmx_net_adam(pmx_net* net, float lr, float beta1, float beta2, float eps) {
    if (!net) return pmx_fail(nullptr, PMX_ERR_ARG, "net is NULL");
    net->adam_t++;
    NCHK(net, pmx_launch_adam(net->ctx->stream, net->master, net->grad, net->adam_m, net->adam_v,
                              net->wb, net->n_total, net->n_w, lr, beta1, beta2, eps,
                              net->adam_t));
    net->wt_fresh = false;
    return PMX_OK;
}

int pmx_net_train_step(pmx_net* net, const void* x, const float* tgt, int N, int H0, int W0,
                       float lr, float w_conf, float* loss) {
    if (!net || !x || !tgt) return pmx_fail(net ? net->ctx : nullptr, PMX_ERR_ARG, "NULL argument");
    int rc = run_forward(net, (const uint16_t*)x, N, H0, W0);
    if (rc == PMX_OK) rc = run_backward(net, (const uint16_t*)x, tgt, w_conf);
    if (rc == PMX_OK) rc = pmx_net_adam(net, lr, 0.9f, 0.999f, 1e-8f);
    if (rc == PMX_OK) rc = copy_sums(net, loss);
    return rc;
}

int pmx_net_propose(pmx_net* net, int W, int H, float tau, float* depth, int depth_pitch,
                    float* spread, int spread_pitch, float* conf, int conf_pitch) {
    if (!net || !depth) return pmx_fail(net ? net->ctx : nullptr, PMX_ERR_ARG, "NULL argument");
    if (net->N != 1)
        return pmx_fail(net->ctx, PMX_ERR_STATE, "propose needs a forward pass of batch 1");
    if ((W + 3) / 4 != net->fW0 || (H + 3) / 4 != net->fH0)
        return pmx_fail(net->ctx, PMX_ERR_ARG, "%dx%d does not match the forward pass (%dx%d)", W,
                        H, net->fW0, net->fH0);
    if (depth_pitch < W * 4 || (conf && conf_pitch < W * 4) || (spread && spread_pitch < W * 4))
        return pmx_fail(net->ctx, PMX_ERR_ARG, "pitch smaller than a row");
    NCHK(net, pmx_launch_proposal(net->ctx->stream, net->head, kHeadC, net->fH0, net->fW0, W, H,
                                  net->inv_lo, net->inv_hi, tau, depth, depth_pitch / 4, spread,
                                  spread_pitch / 4, conf, conf_pitch / 4));
    return PMX_OK;
}

/* ------------------------------------------------------ single operators */

static int conv_args(pmx_ctx* ctx, int N, int H, int W, int Cin, int Cout, int stride) {
    if (!ctx) return pmx_fail(nullptr, PMX_ERR_ARG, "ctx is NULL");
    if (N <= 0 || H <= 0 || W <= 0) return pmx_fail(ctx, PMX_ERR_ARG, "sizes must be positive");
    if (Cin <= 0 || Cout <= 0 || Cin % 32 || Cout % 32)
        return pmx_fail(ctx, PMX_ERR_ARG, "channel counts must be positive multiples of 32");
    if (stride != 1 && stride != 2) return pmx_fail(ctx, PMX_ERR_ARG, "stride must be 1 or 2");
    return PMX_OK;
}

int pmx_cnn_conv_fwd(pmx_ctx* ctx, const void* X, const void* Wt, const float* bias, void* Y,
                     float* Yf, int N, int H, int W, int Cin, int Cout, int stride, int relu) {
    int rc = conv_args(ctx, N, H, W, Cin, Cout, stride);
    if (rc != PMX_OK) return rc;
    if (!X || !Wt || (!Y && !Yf)) return pmx_fail(ctx, PMX_ERR_ARG, "NULL tensor");
    hipError_t e = pmx_launch_conv_fwd(ctx->stream, (const uint16_t*)X, (const uint16_t*)Wt, bias,
                                       (uint16_t*)Y, Yf, N, H, W, Cin, Cout, stride, relu);
    if (e != hipSuccess) return pmx_fail(ctx, PMX_ERR_HIP, "%s", hipGetErrorString(e));
    return PMX_OK;
}

int pmx_cnn_conv_bwd_data(pmx_ctx* ctx, const void* dZ, const void* Wt2, void* dX, int N, int H,
                          int W, int Cin, int Cout, int stride, int accumulate) {
    int rc = conv_args(ctx, N, H, W, Cin, Cout, stride);
    if (rc != PMX_OK) return rc;
    if (!dZ || !Wt2 || !dX) return pmx_fail(ctx, PMX_ERR_ARG, "NULL tensor");
    hipError_t e = pmx_launch_conv_bwd_data(ctx->stream, (const uint16_t*)dZ, (const uint16_t*)Wt2,
                                            (uint16_t*)dX, N, H, W, Cin, Cout, stride, accumulate);
    if (e != hipSuccess) return pmx_fail(ctx, PMX_ERR_HIP, "%s", hipGetErrorString(e));
    return PMX_OK;
}

int pmx_cnn_conv_bwd_weight(pmx_ctx* ctx, const void* X, const void* dZ, float* dW, int N, int H,
                            int W, int Cin, int Cout, int stride) {
    int rc = conv_args(ctx, N, H, W, Cin, Cout, stride);
    if (rc != PMX_OK) return rc;
    if (!X || !dZ || !dW) return pmx_fail(ctx, PMX_ERR_ARG, "NULL tensor");
    hipError_t e = pmx_launch_conv_bwd_weight(ctx->stream, (const uint16_t*)X, (const uint16_t*)dZ,
                                              dW, N, H, W, Cin, Cout, stride);
    if (e != hipSuccess) return pmx_fail(ctx, PMX_ERR_HIP, "%s", hipGetErrorString(e));
    return PMX_OK;
}

int pmx_cnn_weight_t(pmx_ctx* ctx, const void* Wt, void* Wt2, int Cout, int Cin) {
    if (!ctx || !Wt || !Wt2) return pmx_fail(ctx, PMX_ERR_ARG, "NULL argument");
    if (Cout <= 0 || Cin <= 0) return pmx_fail(ctx, PMX_ERR_ARG, "sizes must be positive");
    hipError_t e = pmx_launch_weight_t(ctx->stream, (const uint16_t*)Wt, (uint16_t*)Wt2, Cout, Cin);
    if (e != hipSuccess) return pmx_fail(ctx, PMX_ERR_HIP, "%s", hipGetErrorString(e));
    return PMX_OK;
}

int pmx_cnn_target(pmx_ctx* ctx, const float* depth, int depth_pitch, int W, int H, float* tgt) {
    if (!ctx || !depth || !tgt) return pmx_fail(ctx, PMX_ERR_ARG, "NULL argument");
    if (W <= 0 || H <= 0 || depth_pitch < W * 4)
        return pmx_fail(ctx, PMX_ERR_ARG, "bad size or pitch");
    hipError_t e = pmx_launch_target(ctx->stream, depth, depth_pitch / 4, W, H, (W + 3) / 4,
                                     (H + 3) / 4, tgt);
    if (e != hipSuccess) return pmx_fail(ctx, PMX_ERR_HIP, "%s", hipGetErrorString(e));
    return PMX_OK;
}

}  // extern "C"
