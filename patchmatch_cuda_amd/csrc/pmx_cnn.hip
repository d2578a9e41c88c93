/*
 * pmx_cnn.hip -- the learned depth proposal: a BF16 encoder-decoder CNN with
 * hand-written MFMA convolutions (forward, backward data, backward weights),
 * its loss, Adam, and the glue that turns its output into one proposal depth
 * per pixel for the sweeps.
 *
 * The reference has no learned component; the proposal enters its sweep as a
 * fourth candidate beside depth_array's former / current / random depths
 * (src/patchMatch.cu:827-830, :935-977) and findMinCost stays the judge
 * (src/patchMatch.cu:613-618).
 *
 * Tensors are NHWC: a pixel's channels are contiguous, so one lane's 8
 * consecutive channels are one 16-byte load.  Every channel count is a
 * multiple of 32 (padded with zero channels).  Weights are [Cout][9][Cin]
 * (tap = 3 ky + kx); backward data reads a [Cin][9][Cout] copy.
 *
 * Matrix cores: v_mfma_f32_16x16x32_bf16.  Lane l holds A[l % 16][8 (l / 16)
 * + j] and B[8 (l / 16) + j][l % 16], j = 0..7, and receives D[4 (l / 16) +
 * j][l % 16], j = 0..3.  Accumulation is FP32.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pmx_internal.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float bf2f(uint16_t h) {
    return __uint_as_float((uint32_t)h << 16);
}
/* round to nearest even (NaN kept quiet) */
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *(const bf16x8*)p; }
__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; j++) z[j] = (__bf16)0.0f;
    return z;
}

struct ConvDims {
    int N, H, W, Cin;    /* input */
    int Ho, Wo, Cout;    /* output */
    int stride;
};

/* ------------------------------------------------------------ forward */
/* Y = act(bias + conv3x3(X, Wt)), pad 1.  One wave: 16 output pixels x
 * 16*NT output channels; 4 waves per workgroup over consecutive pixels. */
template <int NT>
__global__ __launch_bounds__(256) void conv_fwd_kernel(const uint16_t* __restrict__ X,
                                                       const uint16_t* __restrict__ Wt,
                                                       const float* __restrict__ bias,
                                                       uint16_t* __restrict__ Y,
                                                       float* __restrict__ Yf, ConvDims d,
                                                       int relu) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = d.N * d.Ho * d.Wo;
    const int p0 = ((int)blockIdx.x * 4 + wave) * 16;
    if (p0 >= P) return;
    const int co0 = (int)blockIdx.y * 16 * NT;
    const int pr = p0 + (lane & 15);
    const bool pin = pr < P;
    int n = 0, oy = 0, ox = 0;
    if (pin) {
        n = pr / (d.Ho * d.Wo);
        const int r = pr - n * d.Ho * d.Wo;
        oy = r / d.Wo;
        ox = r - oy * d.Wo;
    }
    const int kq = (lane >> 4) * 8;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int tap = 0; tap < 9; tap++) {
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
        const int iy = oy * d.stride + ky - 1, ix = ox * d.stride + kx - 1;
        const bool ok = pin && iy >= 0 && iy < d.H && ix >= 0 && ix < d.W;
        const uint16_t* xp = X + (((size_t)n * d.H + (ok ? iy : 0)) * d.W + (ok ? ix : 0)) * d.Cin + kq;
        for (int c0 = 0; c0 < d.Cin; c0 += 32) {
            const bf16x8 a = ok ? ld8(xp + c0) : zero8();
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int co = co0 + t * 16 + (lane & 15);
                const bf16x8 b = co < d.Cout ? ld8(Wt + ((size_t)co * 9 + tap) * d.Cin + c0 + kq)
                                             : zero8();
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co = co0 + t * 16 + (lane & 15);
        if (co >= d.Cout) continue;
        const float bb = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int p = p0 + 4 * (lane >> 4) + j;
            if (p >= P) continue;
            float v = acc[t][j] + bb;
            if (relu) v = fmaxf(v, 0.0f);
            if (Yf) Yf[(size_t)p * d.Cout + co] = v;
            if (Y) Y[(size_t)p * d.Cout + co] = f2bf(v);
        }
    }
}

/* ------------------------------------------------------ backward, data */
/* dX[n][iy][ix][ci] (+)= sum_{tap, co} dZ[n][oy][ox][co] Wt2[ci][tap][co]
 * with oy = (iy + 1 - ky) / stride when that is whole and inside. */
template <int NT>
__global__ __launch_bounds__(256) void conv_bwd_data_kernel(const uint16_t* __restrict__ dZ,
                                                            const uint16_t* __restrict__ Wt2,
                                                            uint16_t* __restrict__ dX,
                                                            ConvDims d, int accumulate) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = d.N * d.H * d.W;
    const int p0 = ((int)blockIdx.x * 4 + wave) * 16;
    if (p0 >= P) return;
    const int ci0 = (int)blockIdx.y * 16 * NT;
    const int pr = p0 + (lane & 15);
    const bool pin = pr < P;
    int n = 0, iy = 0, ix = 0;
    if (pin) {
        n = pr / (d.H * d.W);
        const int r = pr - n * d.H * d.W;
        iy = r / d.W;
        ix = r - iy * d.W;
    }
    const int kq = (lane >> 4) * 8;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int tap = 0; tap < 9; tap++) {
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
        const int ty = iy + 1 - ky, tx = ix + 1 - kx;
        int oy = ty, ox = tx;
        bool ok = pin && ty >= 0 && tx >= 0;
        if (d.stride == 2) {
            ok = ok && !(ty & 1) && !(tx & 1);
            oy = ty >> 1;
            ox = tx >> 1;
        }
        ok = ok && oy < d.Ho && ox < d.Wo;
        const uint16_t* zp = dZ + (((size_t)n * d.Ho + (ok ? oy : 0)) * d.Wo + (ok ? ox : 0)) * d.Cout + kq;
        for (int c0 = 0; c0 < d.Cout; c0 += 32) {
            const bf16x8 a = ok ? ld8(zp + c0) : zero8();
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int ci = ci0 + t * 16 + (lane & 15);
                const bf16x8 b = ci < d.Cin ? ld8(Wt2 + ((size_t)ci * 9 + tap) * d.Cout + c0 + kq)
                                            : zero8();
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int ci = ci0 + t * 16 + (lane & 15);
        if (ci >= d.Cin) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int p = p0 + 4 * (lane >> 4) + j;
            if (p >= P) continue;
            uint16_t* o = dX + (size_t)p * d.Cin + ci;
            float v = acc[t][j];
            if (accumulate) v += bf2f(*o);
            *o = f2bf(v);
        }
    }
}

/* ---------------------------------------------------- backward, weights */
/* dW[co][tap][ci] += sum_p dZ[p][co] X[in(p, tap)][ci] (FP32, atomics over
 * pixel chunks).  A workgroup takes 64 output x 64 input channels of one tap
 * over kChunk output pixels; both operands pass through LDS, transposed so
 * that a lane's 8 consecutive pixels are one 16-byte read. */
constexpr int kChunk = 256;
constexpr int kLdP = 40; /* LDS row: 32 pixels + 8 pad (bf16) */

__global__ __launch_bounds__(256) void conv_bwd_weight_kernel(const uint16_t* __restrict__ X,
                                                              const uint16_t* __restrict__ dZ,
                                                              float* __restrict__ dW, ConvDims d) {
    __shared__ __attribute__((aligned(16))) uint16_t sA[64 * kLdP]; /* [co][px] */
    __shared__ __attribute__((aligned(16))) uint16_t sB[64 * kLdP]; /* [ci][px] */
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int P = d.N * d.Ho * d.Wo;
    const int pbeg = (int)blockIdx.x * kChunk;
    const int co0 = (int)blockIdx.y * 64;
    const int ntci = (d.Cin + 63) / 64;
    const int tap = (int)blockIdx.z / ntci, ci0 = ((int)blockIdx.z % ntci) * 64;
    const int ky = tap / 3, kx = tap - 3 * (tap / 3);
    /* staging: thread -> (pixel tid / 8, channel group tid % 8) */
    const int spx = tid >> 3, scg = (tid & 7) * 8;
    /* compute: wave -> 32 co x 32 ci sub-block */
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kq = (lane >> 4) * 8;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int pk = 0; pk < kChunk; pk += 32) {
        const int p = pbeg + pk + spx;
        /* 8 bf16 values as 4 words; element j is the (j & 1) half of word j / 2 */
        uint4 za = make_uint4(0u, 0u, 0u, 0u), xb = make_uint4(0u, 0u, 0u, 0u);
        if (p < P) {
            const int n = p / (d.Ho * d.Wo);
            const int r = p - n * d.Ho * d.Wo;
            const int oy = r / d.Wo, ox = r - oy * d.Wo;
            if (co0 + scg < d.Cout) za = *(const uint4*)(dZ + (size_t)p * d.Cout + co0 + scg);
            const int iy = oy * d.stride + ky - 1, ix = ox * d.stride + kx - 1;
            if (iy >= 0 && iy < d.H && ix >= 0 && ix < d.W && ci0 + scg < d.Cin)
                xb = *(const uint4*)(X + (((size_t)n * d.H + iy) * d.W + ix) * d.Cin + ci0 + scg);
        }
        __syncthreads(); /* previous step's reads are done */
        const uint32_t zw[4] = {za.x, za.y, za.z, za.w}, xw[4] = {xb.x, xb.y, xb.z, xb.w};
#pragma unroll
        for (int j = 0; j < 8; j++) {
            sA[(scg + j) * kLdP + spx] = (uint16_t)(zw[j >> 1] >> (16 * (j & 1)));
            sB[(scg + j) * kLdP + spx] = (uint16_t)(xw[j >> 1] >> (16 * (j & 1)));
        }
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            a[i] = *(const bf16x8*)(sA + (wm + i * 16 + (lane & 15)) * kLdP + kq);
            b[i] = *(const bf16x8*)(sB + (wn + i * 16 + (lane & 15)) * kLdP + kq);
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int ci = ci0 + wn + j * 16 + (lane & 15);
            if (ci >= d.Cin) continue;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int co = co0 + wm + i * 16 + 4 * (lane >> 4) + e;
                if (co < d.Cout) atomicAdd(&dW[((size_t)co * 9 + tap) * d.Cin + ci], acc[i][j][e]);
            }
        }
}

/* db[c] += sum_p dZ[p][c] */
/* db[c] += sum over pixels of dZ[., c]: each thread sums a strided set of
 * pixels of one channel, the block reduces its rows in LDS, and one atomic
 * per channel per block adds the block's sum (one atomic per thread made
 * 256 K atomics on at most 256 addresses contend). */
__global__ __launch_bounds__(256) void bias_grad_kernel(const uint16_t* __restrict__ dZ, float* db,
                                                        int P, int C) {
    __shared__ float part[256];
    const int c = threadIdx.x % C;
    const int rows = 256 / C;
    const int r0 = threadIdx.x / C;
    float s = 0.0f;
    if (r0 < rows)
        for (int p = (int)blockIdx.x * rows + r0; p < P; p += (int)gridDim.x * rows)
            s += bf2f(dZ[(size_t)p * C + c]);
    part[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < C) {
        float t = 0.0f;
        for (int r = 0; r < rows; r++) t += part[r * C + threadIdx.x];
        atomicAdd(&db[threadIdx.x], t);
    }
}

/* dY *= (Y > 0) */
__global__ void relu_bwd_kernel(uint16_t* __restrict__ dY, const uint16_t* __restrict__ Y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(bf2f(Y[i]) > 0.0f)) dY[i] = 0;
}

/* W [Cout][9][Cin] -> W2 [Cin][9][Cout] */
__global__ void weight_t_kernel(const uint16_t* __restrict__ W, uint16_t* __restrict__ W2, int Cout,
                                int Cin) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= Cout * 9 * Cin) return;
    const int ci = i % Cin, tap = (i / Cin) % 9, co = i / (9 * Cin);
    W2[((size_t)ci * 9 + tap) * Cout + co] = W[i];
}

/* out[n][y][x][0:Ca] = A[n][y/2][x/2][:], out[...][Ca:Ca+Cs] = S[n][y][x][:] */
__global__ void upcat_kernel(const uint16_t* __restrict__ A, const uint16_t* __restrict__ S,
                             uint16_t* __restrict__ out, int N, int H, int W, int Ha, int Wa,
                             int Ca, int Cs) {
    const int C = Ca + Cs;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W * C) return;
    const int c = (int)(i % C);
    const size_t p = i / C;
    const int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((size_t)W * H));
    out[i] = c < Ca ? A[(((size_t)n * Ha + y / 2) * Wa + x / 2) * Ca + c]
                    : S[p * Cs + (c - Ca)];
}

/* backward of upcat: dA[n][ya][xa][c] = sum of the 2x2 block (inside H x W),
 * dS = the skip slice (written, not added) */
__global__ void upcat_bwd_kernel(const uint16_t* __restrict__ dO, uint16_t* __restrict__ dA,
                                 uint16_t* __restrict__ dS, int N, int H, int W, int Ha, int Wa,
                                 int Ca, int Cs) {
    const int C = Ca + Cs;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nA = (size_t)N * Ha * Wa * Ca;
    if (i < nA) {
        const int c = (int)(i % Ca);
        const size_t p = i / Ca;
        const int xa = (int)(p % Wa), ya = (int)((p / Wa) % Ha), n = (int)(p / ((size_t)Wa * Ha));
        float s = 0.0f;
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) {
                const int y = 2 * ya + dy, x = 2 * xa + dx;
                if (y < H && x < W) s += bf2f(dO[(((size_t)n * H + y) * W + x) * C + c]);
            }
        dA[i] = f2bf(s);
    }
    const size_t nS = (size_t)N * H * W * Cs;
    if (i < nS) {
        const int c = (int)(i % Cs);
        const size_t p = i / Cs;
        dS[i] = dO[p * C + Ca + c];
    }
}

/* Head output (FP32, channel 0: inverse-depth logit z0, 1: scale logit z1):
 * inverse depth u = lo + (hi - lo) sigmoid(z0), scale of its error
 * b = kScaleLo + (kScaleHi - kScaleLo) sigmoid(z1).  Loss: L1 on u plus
 * w_conf x L1 of b against the error |u - t| held fixed (b learns the median
 * error; both gradients are bounded).  Gradient into dH (bf16, C channels,
 * zero beyond 1).  tgt <= 0 marks no target. */
constexpr float kScaleLo = 1e-4f, kScaleHi = 0.5f;

__global__ void head_loss_kernel(const float* __restrict__ Hf, const float* __restrict__ tgt,
                                 uint16_t* __restrict__ dH, int P, int C, float lo, float hi,
                                 float inv_count, float w_conf, float* __restrict__ sums) {
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    float l1 = 0.0f, lb = 0.0f;
    if (p < P) {
        const float z0 = Hf[(size_t)p * C], z1 = Hf[(size_t)p * C + 1];
        const float s0 = 1.0f / (1.0f + __expf(-z0));
        const float u = lo + (hi - lo) * s0;
        const float s1 = 1.0f / (1.0f + __expf(-z1));
        const float b = kScaleLo + (kScaleHi - kScaleLo) * s1;
        const float t = tgt[p];
        float g0 = 0.0f, g1 = 0.0f;
        if (t > 0.0f) {
            const float e = u - t;
            l1 = fabsf(e);
            g0 = (e > 0.0f ? 1.0f : (e < 0.0f ? -1.0f : 0.0f)) * (hi - lo) * s0 * (1.0f - s0) * inv_count;
            const float eb = b - fabsf(e);
            lb = fabsf(eb);
            g1 = w_conf * (eb > 0.0f ? 1.0f : (eb < 0.0f ? -1.0f : 0.0f)) * (kScaleHi - kScaleLo) *
                 s1 * (1.0f - s1) * inv_count;
        }
        for (int ch = 0; ch < C; ch++)
            dH[(size_t)p * C + ch] = f2bf(ch == 0 ? g0 : (ch == 1 ? g1 : 0.0f));
    }
    /* block sums */
    for (int o = 32; o >= 1; o >>= 1) {
        l1 += __shfl_xor(l1, o);
        lb += __shfl_xor(lb, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&sums[0], l1);
        atomicAdd(&sums[1], lb);
    }
}

/* Adam (Kingma & Ba 2015) with bias correction; refreshes the bf16 copy of
 * weights (n_w first entries) and clears the gradient. */
__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, uint16_t* __restrict__ wb, size_t n,
                            size_t n_w, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    g[i] = 0.0f;
    if (i < n_w) wb[i] = f2bf(wi);
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ a, uint16_t* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

__global__ void zero_f32_kernel(float* a, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0f;
}

/* Proposal at full resolution: bilinear x4 upsampling (clamp to edge) of the
 * head's inverse depth u and error scale b.  Confidence = 1 - exp(-0.01 u / b),
 * the chance of an error below 1% of u for an exponential error of scale b.  depth = 1/u where the
 * confidence reaches tau, else 0 (no proposal); spread = kSpread b, the
 * half-width in inverse depth of the sweeps' draws around u.  out_spread and
 * out_conf optional. */
constexpr float kSpread = 2.0f;

__global__ void proposal_kernel(const float* __restrict__ Hf, int C, int H0, int W0, int W, int H,
                                float lo, float hi, float tau, float* __restrict__ out, int ld,
                                float* __restrict__ out_spread, int ld_spread,
                                float* __restrict__ out_conf, int ld_conf) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= W || y >= H) return;
    const float fx = fminf(fmaxf((x + 0.5f) * 0.25f - 0.5f, 0.0f), (float)(W0 - 1));
    const float fy = fminf(fmaxf((y + 0.5f) * 0.25f - 0.5f, 0.0f), (float)(H0 - 1));
    const int x0 = (int)fx, y0 = (int)fy;
    const int x1 = min(x0 + 1, W0 - 1), y1 = min(y0 + 1, H0 - 1);
    const float ax = fx - x0, ay = fy - y0;
    float u = 0.0f, c = 0.0f;
    const int xs[2] = {x0, x1}, ys[2] = {y0, y1};
    const float wx[2] = {1.0f - ax, ax}, wy[2] = {1.0f - ay, ay};
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
            const float* h = Hf + ((size_t)ys[j] * W0 + xs[i]) * C;
            const float w = wx[i] * wy[j];
            u += w * (lo + (hi - lo) / (1.0f + __expf(-h[0])));
            c += w * (kScaleLo + (kScaleHi - kScaleLo) / (1.0f + __expf(-h[1])));
        }
    const float b = c;
    const float conf = 1.0f - __expf(-0.01f * u / b);
    out[(size_t)y * ld + x] = (conf >= tau && u > 0.0f) ? 1.0f / u : 0.0f;
    if (out_spread) out_spread[(size_t)y * ld_spread + x] = kSpread * b;
    if (out_conf) out_conf[(size_t)y * ld_conf + x] = conf;
}

/* Training target: mean inverse depth of each 4x4 block of a depth map
 * (0 if any depth is not positive). */
__global__ void target_kernel(const float* __restrict__ depth, int ld, int W, int H, int W0, int H0,
                              float* __restrict__ tgt) {
    const int X = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int Y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (X >= W0 || Y >= H0) return;
    float s = 0.0f;
    int n = 0;
    bool ok = true;
    for (int dy = 0; dy < 4; dy++)
        for (int dx = 0; dx < 4; dx++) {
            const int x = 4 * X + dx, y = 4 * Y + dy;
            if (x >= W || y >= H) continue;
            const float d = depth[(size_t)y * ld + x];
            if (!(d > 0.0f)) ok = false;
            else s += 1.0f / d;
            n++;
        }
    tgt[(size_t)Y * W0 + X] = (ok && n) ? s / n : 0.0f;
}

inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

/* ------------------------------------------------------------ launchers */
hipError_t pmx_launch_conv_fwd(hipStream_t s, const uint16_t* X, const uint16_t* W,
                               const float* bias, uint16_t* Y, float* Yf, int N, int H, int Wd,
                               int Cin, int Cout, int stride, int relu) {
    ConvDims d{N, H, Wd, Cin, (H - 1) / stride + 1, (Wd - 1) / stride + 1, Cout, stride};
    const int P = N * d.Ho * d.Wo;
    const int NT = Cout >= 64 ? 4 : (Cout >= 32 ? 2 : 1);
    dim3 g(nblk(P, 64), (unsigned)((Cout + 16 * NT - 1) / (16 * NT)));
    if (NT == 4) conv_fwd_kernel<4><<<g, 256, 0, s>>>(X, W, bias, Y, Yf, d, relu);
    else if (NT == 2) conv_fwd_kernel<2><<<g, 256, 0, s>>>(X, W, bias, Y, Yf, d, relu);
    else conv_fwd_kernel<1><<<g, 256, 0, s>>>(X, W, bias, Y, Yf, d, relu);
    return hipGetLastError();
}

hipError_t pmx_launch_conv_bwd_data(hipStream_t s, const uint16_t* dZ, const uint16_t* W2,
                                    uint16_t* dX, int N, int H, int Wd, int Cin, int Cout,
                                    int stride, int accumulate) {
    ConvDims d{N, H, Wd, Cin, (H - 1) / stride + 1, (Wd - 1) / stride + 1, Cout, stride};
    const int P = N * H * Wd;
    const int NT = Cin >= 64 ? 4 : (Cin >= 32 ? 2 : 1);
    dim3 g(nblk(P, 64), (unsigned)((Cin + 16 * NT - 1) / (16 * NT)));
    if (NT == 4) conv_bwd_data_kernel<4><<<g, 256, 0, s>>>(dZ, W2, dX, d, accumulate);
    else if (NT == 2) conv_bwd_data_kernel<2><<<g, 256, 0, s>>>(dZ, W2, dX, d, accumulate);
    else conv_bwd_data_kernel<1><<<g, 256, 0, s>>>(dZ, W2, dX, d, accumulate);
    return hipGetLastError();
}

hipError_t pmx_launch_conv_bwd_weight(hipStream_t s, const uint16_t* X, const uint16_t* dZ,
                                      float* dW, int N, int H, int Wd, int Cin, int Cout,
                                      int stride) {
    ConvDims d{N, H, Wd, Cin, (H - 1) / stride + 1, (Wd - 1) / stride + 1, Cout, stride};
    const int P = N * d.Ho * d.Wo;
    dim3 g(nblk(P, kChunk), (unsigned)((Cout + 63) / 64), (unsigned)(9 * ((Cin + 63) / 64)));
    conv_bwd_weight_kernel<<<g, 256, 0, s>>>(X, dZ, dW, d);
    return hipGetLastError();
}

hipError_t pmx_launch_bias_grad(hipStream_t s, const uint16_t* dZ, float* db, int P, int C) {
    if (C < 1 || C > 256) return hipErrorInvalidValue; /* one channel per lane of a block */
    const unsigned blocks = std::min<unsigned>(nblk(P, 256 / C), 512u);
    bias_grad_kernel<<<blocks, 256, 0, s>>>(dZ, db, P, C);
    return hipGetLastError();
}

hipError_t pmx_launch_relu_bwd(hipStream_t s, uint16_t* dY, const uint16_t* Y, size_t n) {
    relu_bwd_kernel<<<nblk(n, 256), 256, 0, s>>>(dY, Y, n);
    return hipGetLastError();
}

hipError_t pmx_launch_weight_t(hipStream_t s, const uint16_t* W, uint16_t* W2, int Cout, int Cin) {
    weight_t_kernel<<<nblk((size_t)Cout * 9 * Cin, 256), 256, 0, s>>>(W, W2, Cout, Cin);
    return hipGetLastError();
}

hipError_t pmx_launch_upcat(hipStream_t s, const uint16_t* A, const uint16_t* S, uint16_t* out,
                            int N, int H, int W, int Ha, int Wa, int Ca, int Cs) {
    upcat_kernel<<<nblk((size_t)N * H * W * (Ca + Cs), 256), 256, 0, s>>>(A, S, out, N, H, W, Ha,
                                                                         Wa, Ca, Cs);
    return hipGetLastError();
}

hipError_t pmx_launch_upcat_bwd(hipStream_t s, const uint16_t* dO, uint16_t* dA, uint16_t* dS,
                                int N, int H, int W, int Ha, int Wa, int Ca, int Cs) {
    const size_t n = std::max((size_t)N * Ha * Wa * Ca, (size_t)N * H * W * Cs);
    upcat_bwd_kernel<<<nblk(n, 256), 256, 0, s>>>(dO, dA, dS, N, H, W, Ha, Wa, Ca, Cs);
    return hipGetLastError();
}

hipError_t pmx_launch_head_loss(hipStream_t s, const float* Hf, const float* tgt, uint16_t* dH,
                                int P, int C, float lo, float hi, float inv_count, float w_conf,
                                float* sums) {
    head_loss_kernel<<<nblk(P, 256), 256, 0, s>>>(Hf, tgt, dH, P, C, lo, hi, inv_count, w_conf,
                                                   sums);
    return hipGetLastError();
}

hipError_t pmx_launch_adam(hipStream_t s, float* w, float* g, float* m, float* v, uint16_t* wb,
                           size_t n, size_t n_w, float lr, float b1, float b2, float eps, int t) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)t));
    const float c2 = 1.0f / (1.0f - powf(b2, (float)t));
    adam_kernel<<<nblk(n, 256), 256, 0, s>>>(w, g, m, v, wb, n, n_w, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t pmx_launch_f32_to_bf16(hipStream_t s, const float* a, uint16_t* b, size_t n) {
    f32_to_bf16_kernel<<<nblk(n, 256), 256, 0, s>>>(a, b, n);
    return hipGetLastError();
}

hipError_t pmx_launch_zero_f32(hipStream_t s, float* a, size_t n) {
    zero_f32_kernel<<<nblk(n, 256), 256, 0, s>>>(a, n);
    return hipGetLastError();
}

hipError_t pmx_launch_proposal(hipStream_t s, const float* Hf, int C, int H0, int W0, int W, int H,
                               float lo, float hi, float tau, float* out, int ld, float* spread,
                               int ld_spread, float* conf, int ld_conf) {
    dim3 b(32, 8), g(nblk(W, 32), nblk(H, 8));
    proposal_kernel<<<g, b, 0, s>>>(Hf, C, H0, W0, W, H, lo, hi, tau, out, ld, spread, ld_spread,
                                    conf, ld_conf);
    return hipGetLastError();
}

hipError_t pmx_launch_target(hipStream_t s, const float* depth, int ld, int W, int H, int W0, int H0,
                             float* tgt) {
    dim3 b(32, 8), g(nblk(W0, 32), nblk(H0, 8));
    target_kernel<<<g, b, 0, s>>>(depth, ld, W, H, W0, H0, tgt);
    return hipGetLastError();
}
