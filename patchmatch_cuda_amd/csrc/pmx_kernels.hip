/*
 * pmx_kernels.hip -- gfx950 kernels of the PatchMatch engine.
 *
 * Hot path: pmx_sweep_kernel (the row-sequential propagation sweep with HMM
 * view selection, src/patchMatch.cu:818-1209) and pmx_cost_all_kernel
 * (src/patchMatch.cu:775-816).  Both evaluate the windowed NCC of
 * src/patchMatch.cu:433-611 with one 16/32/64-lane group per (view, depth)
 * hypothesis: the lanes of a group own one window column each and walk the
 * window rows, so the bilinear taps of neighbouring lanes fall on
 * neighbouring source texels.  A wavefront owns one sweep column and walks
 * its rows; the per-pixel plane-induced homography is composed in FP64 and
 * the window is projected relative to its centre, so FP32 only carries
 * small offsets.  Window sums are accumulated in FP32 around the reference
 * window mean (no cancellation in the variance), and the reference window
 * itself (minus its mean) sits in LDS, prefetched one row ahead.
 *
 * Support kernels: LDS-tiled transposes, xorwow RNG init and depth fill,
 * separable pre-blur, FP64 box sums, FP64 homography composition.
 */
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "pmx_internal.h"

namespace {

constexpr int kLanes = 64;
constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr float kSentinel = 2.0f;       /* src/patchMatch.cu:594-595 */
constexpr float kFix = 0.999f;          /* src/patchMatch.cu:9 */
constexpr float kChange = 1.0f - 0.999f; /* src/patchMatch.cu:10 */
constexpr float kUniformX = 0.83f;      /* src/patchMatch.cu:673 */

/* ------------------------------------------------------ cross-lane ops */
/* DPP controls within a 16-lane row: quad permutations, row mirrors and row
 * shifts.  They run on the VALU (no LDS round trip like ds_bpermute). */
constexpr int kDppQuadXor1 = 0xB1;      /* quad_perm [1,0,3,2] */
constexpr int kDppQuadXor2 = 0x4E;      /* quad_perm [2,3,0,1] */
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowShr = 0x110;       /* + n: row_shr:n */

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}


/* Sum over each 16-lane row, the same value in every lane of the row
 * (each step adds a lane pair symmetrically, so all lanes agree bitwise). */
__device__ __forceinline__ float row_sum16(float v) {
    v += dpp_f<kDppQuadXor1>(v);
    v += dpp_f<kDppQuadXor2>(v);
    v += dpp_f<kDppRowHalfMirror>(v);
    v += dpp_f<kDppRowMirror>(v);
    return v;
}

/* Sums of consecutive 16-lane rows, for values already equal across each
 * row: with two rows per sum rows 1 and 3 receive rows 0+1 and 2+3; with
 * four, row 3 receives the sum of all four (DPP row_bcast:15 and
 * row_bcast:31, a row's lane 15 added to the next row).  Other rows keep
 * their value.  FP32 addition of two terms commutes, so the sums equal an
 * xor butterfly's bit for bit. */
template <int ROWS>
__device__ __forceinline__ float row_group_sum(float v) {
    if (ROWS >= 2)
        v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                   0x142, 0xA, 0xF, false));
    if (ROWS >= 4)
        v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                   0x143, 0x8, 0xF, false));
    return v;
}

/* Inclusive prefix sum over each 16-lane row (FP32). */
__device__ __forceinline__ float row_scan16f(float v) {
    v += dpp_f<kDppRowShr + 1>(v);
    v += dpp_f<kDppRowShr + 2>(v);
    v += dpp_f<kDppRowShr + 4>(v);
    v += dpp_f<kDppRowShr + 8>(v);
    return v;
}

/* Maximum over each 16-lane row, the same in every lane of the row. */
__device__ __forceinline__ float row_max16(float v) {
    v = fmaxf(v, dpp_f<kDppQuadXor1>(v));
    v = fmaxf(v, dpp_f<kDppQuadXor2>(v));
    v = fmaxf(v, dpp_f<kDppRowHalfMirror>(v));
    v = fmaxf(v, dpp_f<kDppRowMirror>(v));
    return v;
}


/* Workgroup barrier that orders LDS only: a wave's LDS writes are complete
 * before any wave passes it, while its global loads stay in flight
 * (__syncthreads also waits for every outstanding global load, which would
 * expose the latency of the next row's prefetch at each barrier). */
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

/* Value of lane `src` (wave-uniform) in every lane. */
__device__ __forceinline__ float lane_f(float v, int src);

/* Sum over lanes 0..n-1 (the other lanes hold 0), the same in every lane. */
__device__ __forceinline__ float wave_sum(float v, int n) {
    const float r = row_sum16(v);
    float s = lane_f(r, 0);
    if (n > 16) s = (s + lane_f(r, 16)) + (lane_f(r, 32) + lane_f(r, 48));
    return s;
}

/* A wave-uniform value in a scalar register (its load is waited for here). */
/* Column-uniform values that only vector code uses, held in vector
 * registers: the scalar registers of the sweep are the scarce ones (the
 * compiler would otherwise keep them there and spill others). */
template <class T>
__device__ __forceinline__ void to_vgpr(T& v) {
    asm volatile("" : "+v"(v));
}
/* A global-memory pointer held in vector registers (the address space is
 * kept, so the accesses stay global loads and stores). */
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> vglobal(T* p) {
    gptr<T> g = (gptr<T>)p;
    asm volatile("" : "+v"(g));
    return g;
}
/* ... or left where the compiler puts it (scalar registers), for kernels
 * whose vector registers are the scarcer ones */
template <bool VEC, class T>
__device__ __forceinline__ gptr<T> col_ptr(T* p) {
    if constexpr (VEC) return vglobal(p);
    return (gptr<T>)p;
}

__device__ __forceinline__ float uniform_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ float lane_f(float v, int src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
}


/* ------------------------------------------------------------------ RNG */
/* Marsaglia xorwow (J. Stat. Softw. 8(14), 2003); seeding by SplitMix64
 * (Steele, Lea, Flood, OOPSLA 2014) of (seed << 32 | id). */
struct Xw {
    uint32_t x, y, z, w, v, d;
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ Xw xw_seed(uint32_t id, uint32_t seed) {
    uint64_t s = ((uint64_t)seed << 32) | (uint64_t)id;
    const uint64_t a = splitmix64(s), b = splitmix64(s), c = splitmix64(s);
    Xw st{(uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32), (uint32_t)c,
          (uint32_t)(c >> 32)};
    if ((st.x | st.y | st.z | st.w | st.v) == 0u) st.x = 1u;
    return st;
}

__device__ __forceinline__ uint32_t xw_next(Xw& s) {
    const uint32_t t = s.x ^ (s.x >> 2);
    s.x = s.y;
    s.y = s.z;
    s.z = s.w;
    s.w = s.v;
    s.v = (s.v ^ (s.v << 4)) ^ (t ^ (t << 1));
    s.d += 362437u;
    return s.d + s.v;
}

/* ((r >> 8) + 1) * 2^-24: exact in FP32, range (0, 1]. */
__device__ __forceinline__ float xw_uniform(Xw& s) {
    return (float)((xw_next(s) >> 8) + 1u) * 5.9604644775390625e-08f;
}

__device__ __forceinline__ Xw xw_load(const pmx_rng_state* p) {
    return Xw{p->x, p->y, p->z, p->w, p->v, p->d};
}

__device__ __forceinline__ void xw_store(pmx_rng_state* p, const Xw& s) {
    p->x = s.x;
    p->y = s.y;
    p->z = s.z;
    p->w = s.w;
    p->v = s.v;
    p->d = s.d;
}

__device__ __forceinline__ Xw xw_uniformize(Xw s) {
    s.x = __builtin_amdgcn_readfirstlane(s.x);
    s.y = __builtin_amdgcn_readfirstlane(s.y);
    s.z = __builtin_amdgcn_readfirstlane(s.z);
    s.w = __builtin_amdgcn_readfirstlane(s.w);
    s.v = __builtin_amdgcn_readfirstlane(s.v);
    s.d = __builtin_amdgcn_readfirstlane(s.d);
    return s;
}

/* block-linear thread id of src/array2d_psng.cu:8-9 */
__device__ __forceinline__ uint32_t seed_id(int x, int y, int width, int bx, int by) {
    const int gx = (width - 1) / bx + 1;
    return (uint32_t)(((y / by) * gx + (x / bx)) * bx * by + (y % by) * bx + (x % bx));
}

__global__ void rng_init_kernel(pmx_rng_state* st, int pitch, int width, int height, int bx,
                                int by, uint32_t seed) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    pmx_rng_state* p = (pmx_rng_state*)((char*)st + (size_t)y * pitch) + x;
    xw_store(p, xw_seed(seed_id(x, y, width, bx, by), seed));
}

/* generate_kernel_float (src/array2d_wrapper.cu:3-20). */
__global__ void fill_depth_kernel(pmx_rng_state* st, int pitch, float* out, int ld, int width,
                                  int height, float near_d, float far_d, int mode) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    pmx_rng_state* p = (pmx_rng_state*)((char*)st + (size_t)y * pitch) + x;
    Xw s = xw_load(p);
    const float u = xw_uniform(s);
    float d;
    if (mode == PMX_INIT_INVERSE) {
        const float lo = 1.0f / far_d, hi = 1.0f / near_d;
        d = 1.0f / ((1.0f - u) * (hi - lo) + lo);
    } else {
        d = (1.0f - u) * (far_d - near_d) + near_d;
    }
    out[(size_t)y * ld + x] = d;
    xw_store(p, s);
}

/* ----------------------------------------------------------- transposes */
/* Plane transpose, 32x32 tile, 32x8 threads; out[x][y] = in[y][x]. */
__global__ __launch_bounds__(256) void transpose_f32_kernel(const float* __restrict__ in,
                                                            float* __restrict__ out, int width,
                                                            int height, int ld_in, int ld_out) {
    __shared__ float tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int x = bx + tx, y = by + ty + k;
        if (x < width && y < height) tile[ty + k][tx] = in[(size_t)y * ld_in + x];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int y = by + tx, x = bx + ty + k; /* out row = x, out column = y */
        if (x < width && y < height) out[(size_t)x * ld_out + y] = tile[tx][ty + k];
    }
}

/* Transpose of E-float records: in[y][x][e] -> out[x][y][e]; T x T records
 * per workgroup (T = 16, or 8 for E > 32 to bound LDS).  Thread t moves the
 * floats f = t + 256 j of every tile row (T E <= 512, so j < 2): the
 * divisions by E that locate them in the transposed tile are done once per
 * thread, not per float. */
__global__ __launch_bounds__(256) void transpose_rec_kernel(const float* __restrict__ in,
                                                            float* __restrict__ out, int width,
                                                            int height, int E, int T, int ld_in,
                                                            int ld_out) {
    extern __shared__ float tile[]; /* T rows x (T*E + 1) */
    const int row_len = T * E + 1;
    const int bx = blockIdx.x * T, by = blockIdx.y * T;
    const int t = threadIdx.x;
    const int nx = min(T, width - bx), ny = min(T, height - by);
    const int span_in = nx * E, span_out = ny * E;
    const float* __restrict__ src = in + (size_t)by * ld_in + (size_t)bx * E;
    float* __restrict__ dst = out + (size_t)bx * ld_out + (size_t)by * E;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int f = t + 256 * j;
        if (f < span_in)
            for (int r = 0; r < ny; r++) tile[r * row_len + f] = src[(size_t)r * ld_in + f];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int f = t + 256 * j;
        if (f < span_out) {
            const int yy = f / E, e = f - yy * E;
            const float* __restrict__ tp = tile + yy * row_len + e;
            for (int xr = 0; xr < nx; xr++) dst[(size_t)xr * ld_out + f] = tp[xr * E];
        }
    }
}

/* Reference layout [v][y][x] <-> engine layout [y][x][v]. */
__global__ void gather_costs_kernel(const float* __restrict__ cost, int cost_ld,
                                    float* __restrict__ out, int out_ld, int width, int height,
                                    int V) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    for (int v = 0; v < V; v++)
        out[((size_t)v * height + y) * out_ld + x] = cost[(size_t)y * cost_ld + (size_t)x * V + v];
}

__global__ void scatter_costs_kernel(const float* __restrict__ in, int in_ld,
                                     float* __restrict__ cost, int cost_ld, int width, int height,
                                     int V) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    for (int v = 0; v < V; v++)
        cost[(size_t)y * cost_ld + (size_t)x * V + v] = in[((size_t)v * height + y) * in_ld + x];
}

/* --------------------------------------------------------- image prep */
struct Taps {
    float t[45];
};

/* Separable filter with clamp-to-edge (src/GaussianBlurCUDA.cu:39-70):
 * row pass into tmp, column pass back. */
__global__ void blur_rows_kernel(const float* __restrict__ in, float* __restrict__ tmp, int ld,
                                 int width, int height, Taps taps, int fr) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    float s = 0.0f;
    for (int k = -fr; k <= fr; k++) {
        const int xx = min(max(x + k, 0), width - 1);
        s = fmaf(in[(size_t)y * ld + xx], taps.t[fr - k], s);
    }
    tmp[(size_t)y * width + x] = s;
}

__global__ void blur_cols_kernel(const float* __restrict__ tmp, float* __restrict__ out, int ld,
                                 int width, int height, Taps taps, int fr) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    float s = 0.0f;
    for (int k = -fr; k <= fr; k++) {
        const int yy = min(max(y + k, 0), height - 1);
        s = fmaf(tmp[(size_t)yy * width + x], taps.t[fr - k], s);
    }
    out[(size_t)y * ld + x] = s;
}

/* Box sums in FP64 (src/array2d_refImg.cu:11-67): row pass, then column
 * pass which also writes the NCC statistics (mean, n*variance clamped at
 * 1e-5 as src/patchMatch.cu:597-600). */
__global__ void box_rows_kernel(const float* __restrict__ img, int ld, int width, int height,
                                int hw, double2* __restrict__ tmp) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    double a = 0.0, b = 0.0;
    const int x0 = max(x - hw, 0), x1 = min(x + hw, width - 1);
    for (int xx = x0; xx <= x1; xx++) {
        const double v = (double)img[(size_t)y * ld + xx];
        a += v;
        b = fma(v, v, b);
    }
    tmp[(size_t)y * width + x] = make_double2(a, b);
}

__global__ void box_cols_kernel(const double2* __restrict__ tmp, const float* __restrict__ img,
                                int img_ld, int width, int height, int hw,
                                float* __restrict__ out_I, float* __restrict__ out_sI,
                                float* __restrict__ out_sII, int out_ld,
                                float2* __restrict__ stats, int stats_ld) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    double a = 0.0, b = 0.0;
    const int y0 = max(y - hw, 0), y1 = min(y + hw, height - 1);
    for (int yy = y0; yy <= y1; yy++) {
        const double2 t = tmp[(size_t)yy * width + x];
        a += t.x;
        b += t.y;
    }
    const size_t o = (size_t)y * out_ld + x;
    if (out_I) out_I[o] = img[(size_t)y * img_ld + x];
    if (out_sI) out_sI[o] = (float)a;
    if (out_sII) out_sII[o] = (float)b;
    if (stats) {
        const int nx = min(x + hw, width - 1) - max(x - hw, 0) + 1;
        const int ny = y1 - y0 + 1;
        const double n = (double)(nx * ny);
        double c1 = b - a * a / n;
        c1 = c1 < 0.00001 ? 0.0 : c1;
        stats[(size_t)y * stats_ld + x] = make_float2((float)(a / n), (float)c1);
    }
}

__global__ void convert_u8_kernel(const uint8_t* __restrict__ in, int in_pitch,
                                  float* __restrict__ out, int out_ld, int width, int height) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= width || y >= height) return;
    out[(size_t)y * out_ld + x] = (float)in[(size_t)y * in_pitch + x] * (1.0f / 255.0f);
}

/* Image::init_relative (src/image.cpp:4-13) in FP64, one thread per view:
 * H1 = Ks Rs Rr^T Kr^-1, H2 = Ks (Rs (Cs - Cr)) n^T Kr^-1, n = (oX, 0, oZ). */
__device__ void m3mul(const double* a, const double* b, double* o) {
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            t[i * 3 + j] = fma(a[i * 3 + 0], b[0 * 3 + j],
                               fma(a[i * 3 + 1], b[1 * 3 + j], a[i * 3 + 2] * b[2 * 3 + j]));
    for (int k = 0; k < 9; k++) o[k] = t[k];
}

__global__ void compose_kernel(const double* __restrict__ cams, int V, double oX, double oZ,
                               double* __restrict__ Hout) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double* Kr = cams;
    const double* Rr = cams + 9;
    const double* Cr = cams + 18;
    const double* Ks = cams + (size_t)(v + 1) * 21;
    const double* Rs = Ks + 9;
    const double* Cs = Ks + 18;
    /* Kr^-1 by cofactors */
    const double a = Kr[0], b = Kr[1], c = Kr[2], d = Kr[3], e = Kr[4], f = Kr[5], g = Kr[6],
                 h = Kr[7], i = Kr[8];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double r = 1.0 / (a * A + b * B + c * C);
    const double Ki[9] = {A * r, -(b * i - c * h) * r, (b * f - c * e) * r,
                          B * r, (a * i - c * g) * r,  -(a * f - c * d) * r,
                          C * r, -(a * h - b * g) * r, (a * e - b * d) * r};
    double RrT[9];
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) RrT[p * 3 + q] = Rr[q * 3 + p];
    double Rrel[9], t9[9];
    m3mul(Rs, RrT, Rrel);
    double tr[3];
    for (int p = 0; p < 3; p++)
        tr[p] = Rs[p * 3] * (Cs[0] - Cr[0]) + Rs[p * 3 + 1] * (Cs[1] - Cr[1]) +
                Rs[p * 3 + 2] * (Cs[2] - Cr[2]);
    double* H1 = Hout + (size_t)v * 18;
    double* H2 = H1 + 9;
    m3mul(Ks, Rrel, t9);
    m3mul(t9, Ki, H1);
    const double n[3] = {oX, 0.0, oZ};
    double tn[9];
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) tn[p * 3 + q] = tr[p] * n[q];
    m3mul(Ks, tn, t9);
    m3mul(t9, Ki, H2);
}

/* ------------------------------------------------------------- NCC core */

__device__ __forceinline__ int xcd_remap(int bid, int nwg) {
    /* Workgroups are dealt round-robin to the 8 XCDs; give each XCD a
     * contiguous run of columns so neighbouring columns share an L2. */
    const int q = nwg >> 3, r = nwg & 7;
    const int xcd = bid & 7, k = bid >> 3;
    return (xcd < r) ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

/* Per-pixel constants shared by every hypothesis of the pixel. */
struct PixelCtx {
    int x, y;            /* original frame */
    int dy_lo, dy_hi;    /* window rows (0..Wn-1) whose reference row is inside */
    bool lane_in;        /* this lane's window column is inside */
    float mu, c1, D;
    int npx;             /* in-bounds window pixels n */
    double inv_n;        /* 1 / n */
    double cx, cy;       /* window centre x + 0.5, y + 0.5 */
    double scale;        /* n . K^-1 (x + 0.5, y + 0.5, 1) */
};

/* The padded source block as one raw buffer resource: 32-bit byte offsets,
 * descriptor in SGPRs.  H16: texels are FP16 (a bilinear row pair is one
 * 4-byte load), else FP32 texel pairs: element x of a row holds t[x] and
 * t[x+1] - t[x] (one 8-byte load, and the row blend is one FMA).  Tex is
 * what a row load returns; it stays raw until the blend, so the conversion
 * does not wait on the load inside the fetch batch. */
template <bool H16>
struct SrcBuf {
    using Tex = typename std::conditional<H16, uint32_t, float2>::type;
    static constexpr uint32_t kShift = H16 ? 1u : 3u; /* log2 bytes per element */
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t layer_bytes, pitch_bytes;
    int srcW, srcH;
};

template <bool H16>
__device__ __forceinline__ SrcBuf<H16> src_buf(const PmxSweepArgs& a) {
    SrcBuf<H16> b;
    /* the record count is a 32-bit field (blocks up to 4 GiB - 1) */
    const uint32_t nrec = (uint32_t)min(a.src_bytes, 0xffffffffL);
    b.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.src_block, (short)0, (int)nrec, 0x00020000);
    b.layer_bytes = (uint32_t)a.src_layer << SrcBuf<H16>::kShift;
    b.pitch_bytes = (uint32_t)a.src_pitch << SrcBuf<H16>::kShift;
    b.srcW = a.srcW;
    b.srcH = a.srcH;
    return b;
}

/*
 * One (view, depth) hypothesis of one pixel.  The plane-induced homography
 * T = H1 - H2 / depth / scale (src/patchMatch.cu:465-473) is formed in FP64
 * and the window is projected relative to its centre c' = T (x+0.5, y+0.5):
 * a sample at offset (dx, dy) lands at
 *     c' + ((ux dx + uy dy), (vx dx + vy dy)) / (1 + cx dx + cy dy),
 * so FP32 only carries offsets of a few pixels.
 */
struct Hyp {
    float ux, uy, vx, vy, cx, cy;
    float fxc, fyc;              /* fractional part of the centre */
    float lox, hix, loy, hiy;    /* relative positions that keep a tap inside the padded block */
    uint32_t base;               /* byte offset of texel floor(c') */
};

/* Bilinear tap with zero border (layered texture with border addressing and
 * linear filtering, src/patchMatch.cu:192-195): position relative to the
 * centre texel, clamped into the 2-pixel zero margin of the block. */
/* floor(x) as an integer in one instruction (v_cvt_flr_i32_f32); |x| is
 * far below 2^31 for every clamped tap position. */
__device__ __forceinline__ int floor_i(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

/* Fractional weights of a tap, kept from its fetch to its blend. */
struct TapW {
    float al, be;
};

/* A tap in two halves, so that many taps' loads can be in flight before the
 * first blend: the two row loads (texel pairs of source rows floor(sy) and
 * floor(sy)+1) ... */
template <class SB>
__device__ __forceinline__ TapW tap_fetch(const SB& sb, const Hyp& q, float sx, float sy,
                                          typename SB::Tex& t0, typename SB::Tex& t1) {
    sx = __builtin_amdgcn_fmed3f(sx, q.lox, q.hix);
    sy = __builtin_amdgcn_fmed3f(sy, q.loy, q.hiy);
    const int ix = floor_i(sx), iy = floor_i(sy);
    /* |iy| < 2^23 and the pitch < 2^23 bytes (source images of at most
     * 2^22 rows and 2^20 - 68 columns): signed 24-bit operands */
    /* iy * pitch + (ix << shift) + base as a shift-add and one signed
     * 24-bit multiply-add (the compiler otherwise re-associates it into
     * three instructions) */
    const uint32_t colb = ((uint32_t)ix << SB::kShift) + q.base;
    uint32_t off;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(off) : "v"(iy), "v"(sb.pitch_bytes), "v"(colb));
    if constexpr (SB::kShift == 1) {
        t0 = __builtin_amdgcn_raw_buffer_load_b32(sb.rsrc, off, 0, 0);
        t1 = __builtin_amdgcn_raw_buffer_load_b32(sb.rsrc, off, sb.pitch_bytes, 0);
    } else {
        t0 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(sb.rsrc, off, 0, 0));
        t1 = __builtin_bit_cast(float2,
                                __builtin_amdgcn_raw_buffer_load_b64(sb.rsrc, off, sb.pitch_bytes, 0));
    }
    /* off stays live past the second load, so the register allocator does
     * not give that load off's register as its destination: a load
     * overwriting the address of the load before it costs a wait state
     * (s_nop) between the two on gfx950, one per tap */
    asm volatile("" ::"v"(off));
    return TapW{__builtin_amdgcn_fractf(sx), __builtin_amdgcn_fractf(sy)};
}

/* Horizontal blend of a texel pair, (1 - al) lo + al hi, in FP32: an FP32
 * pair element holds lo and hi - lo (the difference rounded to FP32 when the
 * block is built, as it would be here), so the blend is lo + al (hi - lo) in
 * one FMA.  FP16 pairs (low half first) are blended as lo + al hi - al lo:
 * two FMAs whose FP16 operands the hardware widens itself (v_fma_mix_f32
 * selects either half of the loaded word), so no conversion instructions
 * are needed. */
__device__ __forceinline__ float row_blend(float al, float2 t) { return fmaf(al, t.y, t.x); }
__device__ __forceinline__ float row_blend(float al, uint32_t w) {
    const float lo = (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
    const float hi = (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
    return fmaf(-al, lo, fmaf(al, hi, lo));
}

/* ... and their bilinear blend, in FP32. */
template <class T>
__device__ __forceinline__ float tap_blend(const TapW& w, T r0, T r1) {
    const float top = row_blend(w.al, r0);
    const float bot = row_blend(w.al, r1);
    return fmaf(w.be, bot - top, top);
}

template <class SB>
__device__ __forceinline__ float tap(const SB& sb, const Hyp& q, float sx, float sy) {
    typename SB::Tex t0, t1;
    const TapW w = tap_fetch(sb, q, sx, sy, t0, t1);
    return tap_blend(w, t0, t1);
}

/* 1/x in FP64: hardware reciprocal refined by two Newton steps (full
 * double precision for normal x). */
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

/* Hm: the V x 18 homographies (H1, H2), a copy in LDS */
template <class SB>
__device__ __forceinline__ Hyp make_hyp(const double* __restrict__ Hm, const SB& sb, int view,
                                        float depth, const PixelCtx& pc) {
    const double* __restrict__ Hv = Hm + view * 18;
    /* the plane homography H1 - H2 / (depth scale), scaled by depth scale
     * (the projection is a ratio: no reciprocal needed) */
    const double sd = (double)depth * pc.scale;
    double h[9];
#pragma unroll
    for (int k = 0; k < 9; k++) h[k] = fma(Hv[k], sd, -Hv[9 + k]);
    const double X0 = fma(h[0], pc.cx, fma(h[1], pc.cy, h[2]));
    const double Y0 = fma(h[3], pc.cx, fma(h[4], pc.cy, h[5]));
    const double Z0 = fma(h[6], pc.cx, fma(h[7], pc.cy, h[8]));
    const double iz = rcp64(Z0);
    const double xc = X0 * iz, yc = Y0 * iz;
    /* centre texel, limited so that every byte offset fits 32 bits; a window
     * whose centre is that far out lies entirely in the zero border */
    const double flx = floor(fmin(fmax(xc, -65536.0), 65536.0));
    const double fly = floor(fmin(fmax(yc, -65536.0), 65536.0));
    const int ixc = (int)flx, iyc = (int)fly;
    Hyp q;
    q.fxc = (float)(xc - flx);
    q.fyc = (float)(yc - fly);
    q.lox = (float)(-PMX_SRC_PAD - ixc);
    q.hix = (float)(sb.srcW - ixc);
    q.loy = (float)(-PMX_SRC_PAD - iyc);
    q.hiy = (float)(sb.srcH - iyc);
    q.ux = (float)((h[0] - xc * h[6]) * iz);
    q.uy = (float)((h[1] - xc * h[7]) * iz);
    q.vx = (float)((h[3] - yc * h[6]) * iz);
    q.vy = (float)((h[4] - yc * h[7]) * iz);
    q.cx = (float)(h[6] * iz);
    q.cy = (float)(h[7] * iz);
    q.base = (uint32_t)view * sb.layer_bytes + (uint32_t)(iyc + PMX_SRC_PAD) * sb.pitch_bytes +
             ((uint32_t)(ixc + PMX_SRC_PAD) << SB::kShift);
    return q;
}

/* A hypothesis' projection as four float4 in LDS (computed once, read by
 * every lane group that sums its window). */
__device__ __forceinline__ void store_hyp(float4* h, const Hyp& q, int step = 1) {
    h[0] = make_float4(q.ux, q.uy, q.vx, q.vy);
    h[step] = make_float4(q.cx, q.cy, q.fxc, q.fyc);
    h[2 * step] = make_float4(q.lox, q.hix, q.loy, q.hiy);
    h[3 * step] = make_float4(__uint_as_float(q.base), 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ Hyp load_hyp(const float4* h) {
    const float4 h0 = h[0], h1 = h[1], h2 = h[2], h3 = h[3];
    Hyp q;
    q.ux = h0.x;
    q.uy = h0.y;
    q.vx = h0.z;
    q.vy = h0.w;
    q.cx = h1.x;
    q.cy = h1.y;
    q.fxc = h1.z;
    q.fyc = h1.w;
    q.lox = h2.x;
    q.hix = h2.y;
    q.loy = h2.z;
    q.hiy = h2.w;
    q.base = __float_as_uint(h3.x);
    return q;
}

/* Window rows [r0, r0 + nr) of one lane's window column; R is the static
 * trip count (0: dynamic).  Sums are taken around K: A = sum e,
 * B = sum e^2, C = sum e R' with e = I' - K.
 * With a static trip count the rows are unrolled without branches, so all
 * taps of the column can be in flight at once: rows i >= nr still fetch a
 * (clamped, in-bounds) tap and read a zero padding row of the LDS window,
 * and their e is selected to 0. */
template <int R, int NB = R, class SB>
__device__ __forceinline__ void accum_rows(const SB& sb, const Hyp& q, int Wn, float dxf,
                                           float dy0, const float* __restrict__ wp, int nr,
                                           bool valid, float& K, float& A, float& B, float& C) {
    const float nxa = fmaf(q.uy, dy0, q.ux * dxf);
    const float nya = fmaf(q.vy, dy0, q.vx * dxf);
    const float dza = fmaf(q.cy, dy0, fmaf(q.cx, dxf, 1.0f));
    auto pos = [&](int i) -> float2 {
        const float fi = (float)i;
        const float rz = __builtin_amdgcn_rcpf(fmaf(q.cy, fi, dza));
        return make_float2(fmaf(fmaf(q.uy, fi, nxa), rz, q.fxc), fmaf(fmaf(q.vy, fi, nya), rz, q.fyc));
    };
    auto sample = [&](int i) -> float {
        const float2 p = pos(i);
        return tap(sb, q, p.x, p.y);
    };
    if constexpr (R > 0) {
        /* the taps' loads in batches of NB rows, then the blends: the
         * scheduling barrier keeps the compiler from sinking each load down
         * to its use, which would make the column R dependent round trips
         * to memory */
        /* the centre sample K goes out with the first batch */
        typename SB::Tex k0, k1;
        TapW kw;
#pragma unroll
        for (int b0 = 0; b0 < R; b0 += NB) {
            constexpr int kB = NB;
            typename SB::Tex t0[kB], t1[kB];
            TapW tw[kB];
            if (b0 == 0) kw = tap_fetch(sb, q, q.fxc, q.fyc, k0, k1);
#pragma unroll
            for (int j = 0; j < kB; j++) {
                if (b0 + j < R) {
                    const float2 p = pos(b0 + j);
                    tw[j] = tap_fetch(sb, q, p.x, p.y, t0[j], t1[j]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            if (b0 == 0) K = tap_blend(kw, k0, k1);
#pragma unroll
            for (int j = 0; j < kB; j++) {
                const int i = b0 + j;
                if (i < R) {
                    const float val = tap_blend(tw[j], t0[j], t1[j]);
                    const float e = (valid && i < nr) ? val - K : 0.0f;
                    A += e;
                    B = fmaf(e, e, B);
                    C = fmaf(e, wp[i * Wn], C);
                }
            }
        }
    } else {
        K = tap(sb, q, q.fxc, q.fyc);
        if (!valid) return;
#pragma unroll 4
        for (int i = 0; i < nr; i++) {
            const float e = sample(i) - K;
            A += e;
            B = fmaf(e, e, B);
            C = fmaf(e, wp[i * Wn], C);
        }
    }
}

/* Value of lane n of each 16-lane row in every lane of the row
 * (DPP row_newbcast:n). */
template <int N>
__device__ __forceinline__ float row_bcast16(float v) {
    return dpp_f<0x150 + N>(v);
}

/* accum_rows for a whole window column (R = Wn rows, r0 = 0) of a pixel
 * whose window rows are all inside the image, with 16-lane groups.  No
 * per-tap masks: lanes whose window column is outside (and the 16th lane)
 * are zeroed by the caller after the loop.  The centring sample K is the
 * window's own centre tap (row R/2 of lane R/2 of the group), broadcast
 * within the group, instead of a separate fetch.  It is fetched with the
 * first batch (so every batch size accumulates the rows in the same order
 * with the same K). */
template <int R, int NB, class SB>
__device__ __forceinline__ void accum_rows_interior(const SB& sb, const Hyp& q, int Wn, float dxf,
                                                    const float* __restrict__ wp, float& K,
                                                    float& A, float& B, float& C) {
    constexpr int HC = R / 2;
    constexpr float dy0 = -(float)HC;
    const float nxa = fmaf(q.uy, dy0, q.ux * dxf);
    const float nya = fmaf(q.vy, dy0, q.vx * dxf);
    const float dza = fmaf(q.cy, dy0, fmaf(q.cx, dxf, 1.0f));
    auto pos = [&](int i) -> float2 {
        const float fi = (float)i;
        const float rz = __builtin_amdgcn_rcpf(fmaf(q.cy, fi, dza));
        return make_float2(fmaf(fmaf(q.uy, fi, nxa), rz, q.fxc), fmaf(fmaf(q.vy, fi, nya), rz, q.fyc));
    };
    typename SB::Tex k0, k1;
    TapW kw;
    float kv = 0.0f; /* this lane's row-HC sample */
#pragma unroll
    for (int b0 = 0; b0 < R; b0 += NB) {
        typename SB::Tex t0[NB], t1[NB];
        TapW tw[NB];
        if (b0 == 0) {
            const float2 p = pos(HC);
            kw = tap_fetch(sb, q, p.x, p.y, k0, k1);
        }
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const int i = b0 + j;
            if (i < R && i != HC) {
                const float2 p = pos(i);
                tw[j] = tap_fetch(sb, q, p.x, p.y, t0[j], t1[j]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (b0 == 0) {
            kv = tap_blend(kw, k0, k1);
            K = row_bcast16<HC>(kv);
        }
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const int i = b0 + j;
            if (i < R) {
                const float val = (i == HC) ? kv : tap_blend(tw[j], t0[j], t1[j]);
                const float e = val - K;
                A += e;
                B = fmaf(e, e, B);
                C = fmaf(e, wp[i * Wn], C);
            }
        }
    }
}

/* 1 - NCC from the centred sums (src/patchMatch.cu:594-610), final
 * normalisation in FP64:  n var' = B - A^2/n,
 * sum I' (Rq - mean) = C + K D  (D = sum (Rq - mean), 0 unless the
 * tile-index shift applies). */
__device__ __forceinline__ float ncc_finish(float A, float B, float C, float K,
                                            const PixelCtx& pc) {
    /* sum I'^2 == 0 <=> every in-bounds sample is 0; the centre sample K is
     * one of them */
    if (K == 0.0f && B == 0.0f) return kSentinel;
    double c2 = (double)B - (double)A * (double)A * pc.inv_n;
    c2 = c2 < 0.00001 ? 0.0 : c2;
    /* 1/sqrt(c1 c2): hardware estimate + two Newton steps (both factors are
     * 0 or >= 1e-5, so p is 0 or a normal number) */
    const double p = (double)pc.c1 * c2;
    if (p == 0.0) return 1.0f;
    double r = __builtin_amdgcn_rsq(p);
    const double hp = 0.5 * p;
    r = r * fma(-hp * r, r, 1.5);
    r = r * fma(-hp * r, r, 1.5);
    return (float)(1.0 - ((double)C + (double)K * (double)pc.D) * r);
}

/* ncc_finish without branches (every lane computes the normalisation; the
 * special cases are selected at the end, with the same values). */
__device__ __forceinline__ float ncc_finish_sel(float A, float B, float C, float K,
                                                const PixelCtx& pc) {
    double c2 = (double)B - (double)A * (double)A * pc.inv_n;
    c2 = c2 < 0.00001 ? 0.0 : c2;
    const double p = (double)pc.c1 * c2;
    double r = __builtin_amdgcn_rsq(p);
    const double hp = 0.5 * p;
    r = r * fma(-hp * r, r, 1.5);
    r = r * fma(-hp * r, r, 1.5);
    const float v = (float)(1.0 - ((double)C + (double)K * (double)pc.D) * r);
    return (K == 0.0f && B == 0.0f) ? kSentinel : (p == 0.0 ? 1.0f : v);
}

/*
 * One pass over nt hypotheses (tview[t], tdep[t]) of one pixel on the NW
 * wavefronts of a workgroup (NW*64/G lane groups).  One group per
 * hypothesis, lanes over window columns; with fewer hypotheses than groups
 * each is split by window rows over k = 2, 4, ... groups so no group idles.
 * If k <= groups per wave, the hypothesis stays inside one wave and its
 * 1 - NCC is written to tout[t]; otherwise each wave writes its partial
 * sums to part[t * NW + slice] and the caller combines them after a
 * barrier (combine_pass).  PRE: the caller passes each lane group's
 * projection in qpre and every group's sums go to part (see below), the
 * other lanes' stores to part[spare + lane].  Returns k.
 */
template <int G, int WN, int NW, int NB = 0, bool INT = false, bool PRE = false, class SB>
__device__ __forceinline__ int eval_pass(const PmxSweepArgs& a, const double* __restrict__ Hm,
                                         const SB& sb, const float* __restrict__ win,
                                         int wave, int lane, int nt, const int* tview,
                                         const float* tdep, float* tout, float4* part,
                                         const PixelCtx& pc, const Hyp* qpre = nullptr,
                                         int spare = 40 * NW) {
    constexpr int NG = kLanes / G;
    constexpr int NGT = NG * NW;
    /* tap rows in flight per lane: a whole column with one wave per
     * workgroup (192 VGPRs, 2 waves per SIMD); 8 rows with more waves,
     * which must fit 512 / (2 NW) VGPRs; NB > 0 sets it */
    constexpr int kNB = NB > 0 ? NB : (NW == 1 ? 16 : 8);
    const int Wn = WN > 0 ? WN : 2 * a.hw + 1;
    /* k = NGT >> ceil(log2(nt)): the largest power of two with k nt <= NGT
     * (1 <= nt <= NGT); shifts instead of divisions by k */
    const int lk = __builtin_ctz((unsigned)NGT) - (nt > 1 ? 32 - __builtin_clz((unsigned)(nt - 1)) : 0);
    const int k = 1 << lk;
    const int gg = wave * NG + lane / G, gl = lane % G;
    const int t = gg >> lk, sub = gg & (k - 1);
    const bool active = t < nt;
    float A = 0.0f, B = 0.0f, C = 0.0f, K = 0.0f;
    if (active) {
        const Hyp q = PRE ? *qpre : make_hyp(Hm, sb, tview[t], tdep[t], pc);
        const int R = (Wn + k - 1) >> lk;
        const int r0 = max(sub * R, pc.dy_lo);
        const int r1 = min(min((sub + 1) * R, Wn), pc.dy_hi + 1);
        const int nr = r1 - r0;
        /* every lane of the group runs the taps (so each holds K); lanes
         * whose window column or rows fall outside the image add nothing */
        const bool valid = pc.lane_in && nr > 0;
        const float dxf = (float)(gl - a.hw);
        const float dy0 = (float)(r0 - a.hw);
        const float* wp = win + r0 * Wn + gl;
        bool done = false;
        if constexpr (INT && G == 16 && WN > 0) {
            /* a whole window column per group, every window row inside the
             * image (wave-uniform: one pixel per wave) */
            if (k == 1 && pc.dy_lo == 0 && pc.dy_hi == Wn - 1) {
                accum_rows_interior<WN, cmin(WN, kNB)>(sb, q, Wn, dxf, win + gl, K, A, B, C);
                if (!pc.lane_in) A = B = C = 0.0f;
                done = true;
            }
        }
        if (done) {
        } else if constexpr (WN > 0) {
            switch (k) {
            case 1: accum_rows<WN, cmin(WN, kNB)>(sb, q, Wn, dxf, dy0, wp, nr, valid, K, A, B, C); break;
            case 2: accum_rows<(WN + 1) / 2, cmin((WN + 1) / 2, kNB)>(sb, q, Wn, dxf, dy0, wp, nr, valid, K, A, B, C); break;
            case 4: accum_rows<(WN + 3) / 4>(sb, q, Wn, dxf, dy0, wp, nr, valid, K, A, B, C); break;
            case 8: accum_rows<(WN + 7) / 8>(sb, q, Wn, dxf, dy0, wp, nr, valid, K, A, B, C); break;
            default: accum_rows<(WN + 15) / 16>(sb, q, Wn, dxf, dy0, wp, nr, valid, K, A, B, C); break;
            }
        } else {
            accum_rows<0>(sb, q, Wn, dxf, dy0, wp, nr, valid, K, A, B, C);
        }
    }
    const int kw = k < NG ? k : NG; /* groups of the hypothesis inside this wave */
    /* the group that ends up holding the hypothesis' sums in this wave */
    int last = 0;
    if constexpr (G == 16) {
        A = row_sum16(A);
        B = row_sum16(B);
        C = row_sum16(C);
        if (kw == 2) {
            A = row_group_sum<2>(A);
            B = row_group_sum<2>(B);
            C = row_group_sum<2>(C);
        } else if (kw == 4) {
            A = row_group_sum<4>(A);
            B = row_group_sum<4>(B);
            C = row_group_sum<4>(C);
        }
        last = kw - 1;
    } else {
#pragma unroll
        for (int m = G / 2; m >= 1; m >>= 1) {
            A += __shfl_xor(A, m, G);
            B += __shfl_xor(B, m, G);
            C += __shfl_xor(C, m, G);
        }
        for (int m = G; m < G * kw; m <<= 1) {
            A += __shfl_xor(A, m, kLanes);
            B += __shfl_xor(B, m, kLanes);
            C += __shfl_xor(C, m, kLanes);
        }
    }
    if constexpr (PRE) {
        /* every lane stores: the group's writer its sums, the others to a
         * spare slot of their own, part[spare + lane] (by default past the
         * at most 24 NW slots of a phase's hypotheses; no branch) */
        const bool wr = active && (k <= NG ? sub == last : (sub & (NG - 1)) == last) && gl == 0;
        part[wr ? t * NW + (k <= NG ? 0 : sub / NG) : spare + lane] = make_float4(A, B, C, K);
    } else if (k <= NG) {
        if (active && sub == last && gl == 0) tout[t] = ncc_finish(A, B, C, K, pc);
    } else if (active && (sub & (NG - 1)) == last && gl == 0) {
        part[t * NW + sub / NG] = make_float4(A, B, C, K);
    }
    return k;
}

/* Second half of a pass whose hypotheses spanned several waves: every wave
 * sums the per-wave partials in slice order and writes the same 1 - NCC. */
template <int NW>
__device__ __forceinline__ void combine_pass(int lane, int nt, int slices, const float4* part,
                                             float* tout, const PixelCtx& pc) {
    if (lane < nt) {
        float4 s = part[lane * NW];
        for (int w = 1; w < slices; w++) {
            const float4 q = part[lane * NW + w];
            s.x += q.x;
            s.y += q.y;
            s.z += q.z;
        }
        tout[lane] = ncc_finish(s.x, s.y, s.z, s.w, pc);
    }
}

/* One lane's share of a reference window, positions idx = lane + 64*k: the
 * loaded values as they arrive, and bit k of `in` (position k inside the
 * image) and of `rd` (inside and read; else the value is 0).  Nothing looks
 * at a value before store_window, so a window prefetched a row ahead is not
 * waited for until it is stored. */
template <int PF>
struct WinVals {
    using Mask = typename std::conditional<(PF <= 32), uint32_t, uint64_t>::type;
    float v[PF];
    Mask in, rd;
};

/* Loads the reference window of pixel (x, y) for one lane's share.
 * qx/qy: tile-index shift of the left (plain) or top (rotated) border
 * (src/patchMatch.cu:536-551). */
template <int PF>
__device__ __forceinline__ void load_window(const PmxSweepArgs& a, int x, int y, int qx, int qy,
                                            int Wn, float inv_wn, int lane, WinVals<PF>& w) {
    /* branch-free: one raw buffer load per position (positions that read
     * nothing load element 0, which is discarded) */
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
        (void*)a.ref, (short)0, a.ref_pitch * a.H * (int)sizeof(float), 0x00020000);
    using Mask = typename WinVals<PF>::Mask;
    w.in = 0;
    w.rd = 0;
#pragma unroll
    for (int k = 0; k < PF; k++) {
        const int idx = lane + kLanes * k;
        const int dyi = (int)(((float)idx + 0.5f) * inv_wn);
        const int dxi = idx - dyi * Wn;
        const int px = x + dxi - a.hw, py = y + dyi - a.hw;
        const bool in = (unsigned)px < (unsigned)a.W && (unsigned)py < (unsigned)a.H;
        const int sx = px - qx, sy = py - qy;
        const bool rd = in && (sx | sy) >= 0 && idx < Wn * Wn;
        const uint32_t off = rd ? (uint32_t)(sy * a.ref_pitch + sx) * 4u : 0u;
        w.v[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rr, off, 0, 0));
        w.in |= (Mask)(in ? 1 : 0) << k;
        w.rd |= (Mask)(rd ? 1 : 0) << k;
    }
}

/* The reference windows along one sweep column, one lane's share: position
 * k of the window at sweep row r lies at varying coordinate r + dv[k] (y in
 * a plain sweep, x in a rotated one) and at a fixed coordinate along the
 * other axis, which carries the column's tile-index shift.  Its byte offset
 * in the reference is r * rstride + off[k]; positions outside the image
 * along the varying axis fall outside the buffer or are discarded by their
 * flags, and a raw buffer load outside the buffer returns 0 without a
 * fault.  fix_in / fix_rd hold the positions' flags as if the varying
 * coordinate were inside.  For the small windows (a few positions per
 * lane) whose per-lane constants fit the registers. */
template <int PF>
struct ColWindow {
    static_assert(PF <= 8, "column window constants for small windows only");
    int dv[PF], off[PF];
    uint32_t fix_in, fix_rd;
    int rstride, Lv;
    __amdgpu_buffer_rsrc_t rsrc;
};

template <int PF>
__device__ __forceinline__ ColWindow<PF> col_window(const PmxSweepArgs& a, bool rot, int c, int qf,
                                                   int Wn, float inv_wn, int lane) {
    ColWindow<PF> cw;
    const int Lf = rot ? a.H : a.W;
    cw.fix_in = 0u;
    cw.fix_rd = 0u;
#pragma unroll
    for (int k = 0; k < PF; k++) {
        const int idx = lane + kLanes * k;
        const int dyi = (int)(((float)idx + 0.5f) * inv_wn);
        const int dxi = idx - dyi * Wn;
        const int dv = (rot ? dxi : dyi) - a.hw;
        const int bf = c + (rot ? dyi : dxi) - a.hw; /* fixed coordinate */
        /* positions past the window hold 0 (the zero rows after it) */
        const bool inF = (unsigned)bf < (unsigned)Lf && idx < Wn * Wn;
        const bool rdF = inF && bf - qf >= 0 && idx < Wn * Wn;
        cw.dv[k] = dv;
        cw.off[k] = 4 * (rot ? (bf - qf) * a.ref_pitch + dv : dv * a.ref_pitch + (bf - qf));
        cw.fix_in |= (inF ? 1u : 0u) << k;
        cw.fix_rd |= (rdF ? 1u : 0u) << k;
    }
    cw.rstride = rot ? 4 : 4 * a.ref_pitch;
    cw.Lv = rot ? a.W : a.H;
    cw.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.ref, (short)0,
                                                a.ref_pitch * a.H * (int)sizeof(float), 0x00020000);
    return cw;
}

/* The window of sweep row r of the column. */
template <int PF>
__device__ __forceinline__ void load_col_window(const ColWindow<PF>& cw, int r, WinVals<PF>& w) {
    const uint32_t rb = (uint32_t)(r * cw.rstride);
    uint32_t inv = 0u;
#pragma unroll
    for (int k = 0; k < PF; k++) {
        const bool inV = (unsigned)(r + cw.dv[k]) < (unsigned)cw.Lv;
        w.v[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                               cw.rsrc, rb + (uint32_t)cw.off[k], 0, 0));
        inv |= (inV ? 1u : 0u) << k;
    }
    w.in = cw.fix_in & inv;
    w.rd = cw.fix_rd & inv;
}

/* Writes a loaded window as R' = Rq - mean into LDS (0 outside the image).
 * With need_d (the tile-index shift applies) returns the wave sum D of R'
 * over the in-bounds positions; otherwise D is exactly 0 (the in-bounds
 * values sum to the box sum) and 0 is returned.  ALL: every position is
 * written, those past the window with 0 (their flags are clear), which the
 * caller's buffer must hold (64 PF <= 2 Wn^2, window and zero rows). */
template <int PF, bool ALL = false>
__device__ __forceinline__ float store_window(float* __restrict__ win, const WinVals<PF>& w,
                                              float mu, int Wn, int lane, bool need_d) {
    float d = 0.0f;
#pragma unroll
    for (int k = 0; k < PF; k++) {
        const int idx = lane + kLanes * k;
        if (ALL || idx < Wn * Wn) {
            const float v = ((w.rd >> k) & 1u) ? w.v[k] : 0.0f;
            const float r = ((w.in >> k) & 1u) ? v - mu : 0.0f;
            win[idx] = r;
            d += r;
        }
    }
    if (!need_d) return 0.0f;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) d += __shfl_xor(d, m, kLanes);
    return d;
}

__device__ __forceinline__ void pixel_ctx(const PmxSweepArgs& a, int x, int y, int gl,
                                          PixelCtx& pc) {
    const int hw = a.hw, Wn = 2 * hw + 1;
    pc.x = x;
    pc.y = y;
    pc.dy_lo = max(0, hw - y);
    pc.dy_hi = min(Wn - 1, a.H - 1 - y + hw);
    const int px = x + gl - hw;
    pc.lane_in = (gl < Wn) && px >= 0 && px < a.W;
    const int nx = min(x + hw, a.W - 1) - max(x - hw, 0) + 1;
    const int ny = pc.dy_hi - pc.dy_lo + 1;
    pc.npx = nx * ny;
    pc.inv_n = rcp64((double)pc.npx);
    pc.cx = (double)x + 0.5;
    pc.cy = (double)y + 0.5;
    pc.scale = a.oX * (a.iK0 * pc.cx + a.iK1) + a.oZ;
}

/* ------------------------------------------------------ cost all views */
/* Copies the V x 18 homographies into LDS (read by every hypothesis). */
__device__ __forceinline__ void load_homographies(const PmxSweepArgs& a, double* Hs) {
    for (int i = (int)threadIdx.x; i < a.V * 18; i += (int)blockDim.x) Hs[i] = a.Hm[i];
}

/* Pixel tiles of the cost-all launch: kCostTW x kCostTH pixels per tile,
 * tiles row-major over the image, pixels row-major inside a tile. */
constexpr int kCostTW = 64, kCostTH = 16;

/* computeAllCostGivenDepth (src/patchMatch.cu:775-816): one wavefront per
 * pixel, G-lane groups over views.  Writes the row-major state.  Workgroups
 * walk the image in 64 x 16 pixel tiles, and each XCD takes a contiguous
 * run of tiles: the windows of a tile's 16 rows overlap in the sources, so
 * an XCD's L2 serves most taps (in row-major order the 15-row reuse spans
 * whole image rows of every view, more than the L2 holds). */
template <int G, int PF, int WN, bool H16>
__global__ __launch_bounds__(64) void pmx_cost_all_kernel(PmxSweepArgs a) {
    const int pid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int tiles_x = (a.W + kCostTW - 1) / kCostTW;
    const int tile = pid / (kCostTW * kCostTH), in_tile = pid - tile * (kCostTW * kCostTH);
    const int x = (tile % tiles_x) * kCostTW + in_tile % kCostTW;
    const int y = (tile / tiles_x) * kCostTH + in_tile / kCostTW;
    /* the whole workgroup is one pixel: it leaves together, before any barrier */
    if (x >= a.W || y >= a.H) return;
    extern __shared__ double smem_d[];
    const int Wn = WN > 0 ? WN : 2 * a.hw + 1;
    double* Hs = smem_d;                               /* V x 18 */
    float4* part = (float4*)(smem_d + a.V * 18);       /* 128: sums per view, then spare */
    float4* hypb = part + 128;                         /* V projections, 4 float4 each */
    float* win = (float*)(hypb + 4 * a.V);             /* Wn rows + Wn zero rows, x Wn */
    const int lane = threadIdx.x;
    const int gl = lane % G;
    constexpr int NG = kLanes / G;
    for (int i = Wn * Wn + lane; i < 2 * Wn * Wn; i += kLanes) win[i] = 0.0f;
    load_homographies(a, Hs);
    const float inv_wn = 1.0f / (float)Wn;
    const SrcBuf<H16> sb = src_buf<H16>(a);
    PixelCtx pc;
    pixel_ctx(a, x, y, gl, pc);
    const float2 st = a.stats[(size_t)y * a.stats_pitch + x];
    pc.mu = st.x;
    pc.c1 = st.y;
    /* computeAllCostGivenDepth runs in the rotated frame: the tile shift is
     * along y (src/patchMatch.cu:794). */
    const int qx = a.rotated ? 0 : max(0, a.hw - x);
    const int qy = a.rotated ? max(0, a.hw - y) : 0;
    WinVals<PF> vals;
    load_window<PF>(a, x, y, qx, qy, Wn, inv_wn, lane, vals);
    pc.D = store_window<PF>(win, vals, pc.mu, Wn, lane, qx != 0 || qy != 0);
    /* state is row-major: column c = y, row r = x */
    const float depth = a.depth[(size_t)y * a.depth_ld + x];
    /* each view's projection once, by lane v; the passes leave their window
     * sums in part[v] (spare stores at part[64 + lane]); lane v forms view
     * v's 1 - NCC */
    if (lane < a.V) store_hyp(hypb + 4 * lane, make_hyp(Hs, sb, lane, depth, pc));
    __syncthreads();
    const int gg = lane / G;
    for (int v0 = 0; v0 < a.V; v0 += NG) {
        const int ntp = min(NG, a.V - v0);
        const int lk = __builtin_ctz((unsigned)NG) - (ntp > 1 ? 32 - __builtin_clz((unsigned)(ntp - 1)) : 0);
        const Hyp q = load_hyp(hypb + 4 * (v0 + min(gg >> lk, ntp - 1)));
        eval_pass<G, WN, 1, 0, false, true>(a, Hs, sb, win, 0, lane, ntp, nullptr, nullptr, nullptr,
                                            part + v0, pc, &q, 64 - v0);
    }
    __syncthreads();
    float* cst = a.cost + (size_t)y * a.cost_ld + (size_t)x * a.V;
    if (lane < a.V) {
        const float4 s4 = part[lane];
        cst[lane] = ncc_finish(s4.x, s4.y, s4.z, s4.w, pc);
    }
    /* the evaluation counts of this kernel are a closed form of the image
     * size (counted on the host): no per-pixel atomics on shared counters */
}

/* Plane-sweep cost volume for the proposal network: one wavefront per
 * 1/4-resolution pixel (X, Y), whose window sits at full-resolution pixel
 * (min(4X+2, W-1), min(4Y+2, H-1)); its 64 hypotheses are 32 fronto-parallel
 * planes uniform in inverse depth on two source views.  Output: bf16 NHWC,
 * 64 channels: mean 1 - NCC of the two views per plane (clamped to [0, 2]),
 * the reference intensity, zeros. */
template <int G, int PF, int WN, bool H16>
__global__ __launch_bounds__(64) void pmx_volume_kernel(PmxSweepArgs a, int view0, int view1,
                                                       uint16_t* __restrict__ out, int W0, int H0) {
    extern __shared__ double smem_d[];
    const int Wn = WN > 0 ? WN : 2 * a.hw + 1;
    double* Hs = smem_d;                               /* V x 18 */
    float4* part = (float4*)(smem_d + a.V * 18);       /* 128: sums per hypothesis, then spare */
    float4* hypb = part + 128;                         /* 64 projections, 4 float4 each */
    float* tout = (float*)(hypb + 4 * 64);             /* 64 */
    float* win = tout + 64;                            /* Wn rows + Wn zero rows, x Wn */
    const int lane = threadIdx.x;
    const int gl = lane % G;
    constexpr int NG = kLanes / G;
    for (int i = Wn * Wn + lane; i < 2 * Wn * Wn; i += kLanes) win[i] = 0.0f;
    load_homographies(a, Hs);
    const int pid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int X = pid % W0, Y = pid / W0;
    const int x = min(4 * X + 2, a.W - 1), y = min(4 * Y + 2, a.H - 1);
    const float inv_wn = 1.0f / (float)Wn;
    const SrcBuf<H16> sb = src_buf<H16>(a);
    PixelCtx pc;
    pixel_ctx(a, x, y, gl, pc);
    const float2 st = a.stats[(size_t)y * a.stats_pitch + x];
    pc.mu = st.x;
    pc.c1 = st.y;
    const int qx = max(0, a.hw - x);
    WinVals<PF> vals;
    load_window<PF>(a, x, y, qx, 0, Wn, inv_wn, lane, vals);
    pc.D = store_window<PF>(win, vals, pc.mu, Wn, lane, qx != 0);
    {
        /* hypothesis `lane`: plane lane & 31 on view0 (lanes 0-31) or view1,
         * projected once by its lane; the passes leave the window sums in
         * part, each lane forms its hypothesis' 1 - NCC */
        const int k = lane & 31;
        const float ulo = 1.0f / a.far_d, uhi = 1.0f / a.near_d;
        const float dk = 1.0f / (ulo + (uhi - ulo) * ((float)k + 0.5f) * (1.0f / 32.0f));
        store_hyp(hypb + 4 * lane, make_hyp(Hs, sb, lane < 32 ? view0 : view1, dk, pc));
    }
    __syncthreads();
    for (int t0 = 0; t0 < 64; t0 += NG) {
        const Hyp q = load_hyp(hypb + 4 * (t0 + lane / G));
        eval_pass<G, WN, 1, 0, false, true>(a, Hs, sb, win, 0, lane, NG, nullptr, nullptr, nullptr,
                                            part + t0, pc, &q, 64 - t0);
    }
    __syncthreads();
    {
        const float4 s4 = part[lane];
        tout[lane] = ncc_finish(s4.x, s4.y, s4.z, s4.w, pc);
    }
    __syncthreads();
    float v = 0.0f;
    if (lane < 32)
        v = 0.5f * (fminf(fmaxf(tout[lane], 0.0f), 2.0f) + fminf(fmaxf(tout[32 + lane], 0.0f), 2.0f));
    else if (lane == 32)
        v = a.ref[(size_t)y * a.ref_pitch + x];
    /* bf16, round to nearest even */
    const uint32_t u = __float_as_uint(v);
    out[((size_t)Y * W0 + X) * 64 + lane] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

/* FP16 storage: rounds n floats to the nearest FP16 value (ties to even)
 * in place and, when out is not null, stores the FP16 bits there. */
/* FP32 texel pairs (t[i], t[i+1] - t[i]) of a padded block of n texels;
 * the last texel's neighbour is 0 (padding). */
__global__ void pair_f32_kernel(const float* __restrict__ data, float2* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float t = data[i];
    const float u = i + 1 < n ? data[i + 1] : 0.0f;
    out[i] = make_float2(t, u - t);
}

__global__ void quantize_f16_kernel(float* __restrict__ data, uint16_t* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const _Float16 h = (_Float16)data[i];
    data[i] = (float)h;
    if (out) out[i] = __builtin_bit_cast(uint16_t, h);
}

/* depth = prop where prop > 0 (both in the same layout) */
__global__ void seed_depth_kernel(const float* __restrict__ prop, int prop_ld,
                                  float* __restrict__ depth, int depth_ld, int width, int height) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= width || y >= height) return;
    const float p = prop[(size_t)y * prop_ld + x];
    if (p > 0.0f) depth[(size_t)y * depth_ld + x] = p;
}

/* ----------------------------------------------------------- messages */
/* The two-state view-selection HMM (src/patchMatch.cu:669-715) carried as
 * log-odds l = log(m / (1 - m)) of "view visible", in FP32.  With the
 * emission e(c) = exp(-c^2 / 2 alpha^2) of a visible view and the uniform
 * level u = e(0.83), the reference's updates are
 *     forward   m' = (m fix + (1-m) chg) e / [(m fix + (1-m) chg) e + (m chg + (1-m) fix) u]
 *     backward  m' = (m e fix + (1-m) u chg) / [(m e fix + (1-m) u chg) + (m e chg + (1-m) u fix)]
 * which in log-odds are  l' = G(l) + E  and  l' = G(l + E),  with the
 * emission's log ratio E = log(e / u) = (0.83^2 - c^2) / (2 alpha^2) and the
 * transition G(l) = log((fix e^l + chg) / (chg e^l + fix)), an odd function
 * bounded by log(fix / chg).  The posterior of view v at a row is
 * sigmoid(l_forward + l_backward).  Nothing underflows: the costs that the
 * left/top-border tile-index quirk produces (|c| up to ~100) give E ~ -10^4,
 * representable, where probabilities would be 0.  Constants keep their
 * FP32 literal values. */
struct MsgConst {
    float inv2a2; /* 0.5 / alpha^2 */
    float ex;     /* 0.5 x^2 / alpha^2, x = 0.83f */
};

__device__ __forceinline__ MsgConst msg_const(double alpha2) {
    MsgConst mc;
    const double x = (double)kUniformX;
    mc.inv2a2 = (float)(0.5 / alpha2);
    mc.ex = (float)(0.5 * x * x / alpha2);
    return mc;
}

/* Lanes whose cost is the sentinel, among the view lanes vmask (the
 * compare's own lane mask: a ballot of the bool would select and compare it
 * again). */
__device__ __forceinline__ uint64_t sentinel_lanes(float cost, uint64_t vmask) {
    return __builtin_amdgcn_fcmpf(cost, kSentinel, 1 /* FCMP_OEQ */) & vmask;
}

/* Sentinel draws: every view whose stored cost is exactly 2.0 consumes one
 * uniform, in view order, and uses 2u as its cost (src/patchMatch.cu:680-682,
 * :706-707).  vmask: the lanes of the views. */
__device__ __forceinline__ float sentinel_emission(float cost, bool is_view, int lane, Xw& rng,
                                                   uint64_t vmask) {
    const bool sent = is_view && (cost == kSentinel);
    const uint64_t mask = sentinel_lanes(cost, vmask);
    const int k = __popcll(mask);
    float e = cost;
    if (k) {
        const int rank = __popcll(mask & ((1ull << lane) - 1ull));
        for (int i = 0; i < k; i++) {
            const float u = xw_uniform(rng);
            if (sent && rank == i) e = u * 2.0f;
        }
    }
    return e;
}

/* E = log(e(c) / u) */
__device__ __forceinline__ float emission_lr(float c, const MsgConst& mc) {
    return fmaf(-(c * c), mc.inv2a2, mc.ex);
}

/* G(l) = sign(l) log((fix + chg t) / (chg + fix t)), t = e^-|l| (no overflow) */
__device__ __forceinline__ float trans_g(float l) {
    const float t = __expf(-fabsf(l));
    /* both arguments lie in [min(fix, chg), fix + chg]: the hardware log2
     * (no denormal handling needed), one scaling of the difference */
    const float g = (__builtin_amdgcn_logf(fmaf(kChange, t, kFix)) -
                     __builtin_amdgcn_logf(fmaf(kFix, t, kChange))) * 0.693147180559945309f;
    return copysignf(g, l);
}

/* ---------------------------------------------------------------- sweep */
/* Hypotheses of one row's phase B: up to 3 per sampled view (former,
 * random, proposal), at most 64 sampled views. */
constexpr int kTasks = 3 * 64;
/* Hypotheses of one phase in the small-view kernels: at most 3 per sampled
 * view in phase B (at most 7 draws), at most 16 in phase C. */
constexpr int kHypMax = 24;

/* lowbias32 integer hash (C. Wellons, public domain constants). */
__device__ __forceinline__ uint32_t mix32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

/* The proposal's hypothesis of one pixel in one sweep: inverse depth
 * 1/depth + spread * h, h uniform in [-1, 1) from a hash of the pixel and
 * the sweep (the column's generator is not touched), clamped to
 * [1/far, 1/near].  FP32 with correctly rounded divisions, so the oracle
 * reproduces it bit for bit. */
__device__ __forceinline__ float proposal_draw(float depth, float spread, int x, int y,
                                               const PmxSweepArgs& a) {
    const uint32_t k = mix32((uint32_t)(y * a.W + x) ^
                             mix32((uint32_t)a.sweep_id * 0x9E3779B9u + 0x632BE5ABu));
    const float h = (float)(k >> 8) * (1.0f / 8388608.0f) - 1.0f;
    const float u = fminf(fmaxf(fmaf(spread, h, 1.0f / depth), 1.0f / a.far_d), 1.0f / a.near_d);
    return 1.0f / u;
}

/*
 * topToDown / downToTop (src/patchMatch.cu:818-1209).  One workgroup of NW
 * wavefronts per sweep column.  Every wave runs the column's sequential
 * logic (HMM messages, posterior, draws, argmin) on identical data, so no
 * barrier is needed for it; lanes 0..V-1 carry the per-view state.  For
 * each row the sampled views are deduplicated and the (view, depth)
 * hypotheses are spread over all NW*64/G lane groups of the workgroup.  The
 * NCC of a view at the winning depth is reused when the view was sampled
 * (the reference recomputes the same value).  Wave 0 alone writes the
 * column's outputs.
 */
template <int G, int PF, int WN, int NW, bool H16, int NB = 0, bool SM = false, int ROT = -1,
          int PRP = -1>
__device__ __forceinline__ void sweep_column(const PmxSweepArgs& a, const int c) {
    /* sweep frame and proposal presence, fixed at compile time when ROT / PRP
     * are 0 or 1 (-1: from the arguments) */
    const bool rot = ROT >= 0 ? ROT == 1 : a.rotated != 0;
    const bool prp = PRP >= 0 ? PRP == 1 : a.prop != nullptr;
    extern __shared__ double smem_d[];
    const int Wn = WN > 0 ? WN : 2 * a.hw + 1;
    const int WW = 2 * Wn * Wn;                          /* window + zero rows */
    double* Hs = smem_d;                                 /* V x 18 */
    float4* part = (float4*)(smem_d + a.V * 18);         /* 2 x 64 x NW partials */
    float* winbuf = (float*)(part + 2 * 64 * NW);        /* 2 x WW */
    float* tres = winbuf + 2 * WW;                       /* task results: kTasks, then 64 */
    int* tview = (int*)(tres + kTasks + 64);             /* task views, kTasks */
    float* tdep = (float*)(tview + kTasks);              /* task depths, kTasks */
    float4* hypb = (float4*)(tdep + kTasks);             /* kHypMax projections, 4 float4 each */

    constexpr int NG = kLanes / G;
    constexpr int NGT = NG * NW;
    const int wave = (int)threadIdx.x / kLanes, lane = (int)threadIdx.x % kLanes;
    const bool lead = wave == 0;
    const int gl = lane % G;
    for (int i = (int)threadIdx.x; i < 2 * WW; i += 64 * NW) winbuf[i] = 0.0f;
    load_homographies(a, Hs);
    const int V = a.V, Hs_rows = a.Hs;
    const bool is_view = lane < V;
    const int vl = is_view ? lane : 0;
    const uint64_t vmask_v = (V >= 64) ? ~0ull : ((1ull << V) - 1ull); /* the view lanes */
    const float inv_wn = 1.0f / (float)Wn;
    const SrcBuf<H16> sb = src_buf<H16>(a);
    const MsgConst mc = msg_const(a.alpha2);

    /* this column's depth, costs, backward messages, proposal and window
     * statistics (x = c in a plain sweep, y = c in a rotated one) */
    /* Kernels with a third resident wave per SIMD (NB > 0: tap batches cut
     * to fit 168 registers) keep the column pointers in scalar registers
     * and project phase C's hypotheses again instead of holding phase B's */
    constexpr bool kVec = NB == 0;
    const gptr<float> dcol = col_ptr<kVec>(a.depth + (size_t)c * a.depth_ld);
    const gptr<float> ccol = col_ptr<kVec>(a.cost + (size_t)c * a.cost_ld);
    const gptr<float> scol = col_ptr<kVec>(a.spm + (size_t)c * a.spm_ld);
    const gptr<const float> pcol = col_ptr<kVec>(prp ? a.prop + (size_t)c * a.prop_ld : a.depth);
    const gptr<const float> qcol = col_ptr<kVec>(prp ? a.prop_spread + (size_t)c * a.prop_ld : a.depth);
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    const gptr<const f32x2> stcol =
        col_ptr<kVec>((const f32x2*)(a.stats + (rot ? (size_t)c * a.stats_pitch : (size_t)c)));
    const int ststep = rot ? 1 : a.stats_pitch;

    /* generator of this column: row 0 (plain) or column 0 (rotated) */
    pmx_rng_state* rp = rot ? (pmx_rng_state*)((char*)a.rng + (size_t)c * a.rng_pitch)
                                  : a.rng + c;
    Xw rng = xw_uniformize(xw_load(rp));

    /* ---- backward messages (src/patchMatch.cu:858-869, :1050-1061), as
     * log-odds: m = 0.5 is l = 0 ---- */
    float m = 0.0f;
    {
        int rb = a.down ? Hs_rows - 1 : 0;
        const int sbs = a.down ? -1 : 1;
        float cn = ccol[(size_t)rb * V + vl];
        for (int k = 0; k < Hs_rows; k++, rb += sbs) {
            const float cst = cn;
            if (k + 1 < Hs_rows) cn = ccol[(size_t)(rb + sbs) * V + vl];
            if (lead && is_view) scol[(size_t)rb * V + lane] = m;
            const float e = sentinel_emission(cst, is_view, lane, rng, vmask_v);
            m = trans_g(m + emission_lr(e, mc));
        }
    }
    /* the backward messages are re-read below by every wave */
    __threadfence_block();
    __syncthreads();
    const int step = a.down ? 1 : -1;
    const int r0 = a.down ? 0 : Hs_rows - 1;
    /* original-frame coordinates of sweep row r in this column */
    auto orig_xy = [&](int r, int& x, int& y) {
        if (rot) {
            x = r;
            y = c;
        } else {
            x = c;
            y = r;
        }
    };
    /* ---- forward message at the first row ---- */
    {
        const float cst = ccol[(size_t)r0 * V + vl];
        const float e = sentinel_emission(cst, is_view, lane, rng, vmask_v);
        m = emission_lr(e, mc); /* G(0) = 0 */
    }
    float d_former = uniform_f(dcol[r0]);

    const int qx_col = rot ? 0 : max(0, a.hw - c);
    const int qy_col = rot ? max(0, a.hw - c) : 0;
    const bool need_d = qx_col != 0 || qy_col != 0;

    /* prefetch row r0 + step */
    int r = r0 + step;
    float pf_cost = 0.0f, pf_depth = 0.0f, pf_prop = 0.0f, pf_spread = 0.0f;
    float pf_beta = 0.0f;
    float2 pf_stat = make_float2(0.0f, 0.0f);
    WinVals<PF> pf_win;
    constexpr int kColPF = PF <= 8 ? PF : 1; /* column constants for small windows only */
    const ColWindow<kColPF> cwin =
        col_window<kColPF>(a, rot, c, rot ? qy_col : qx_col, Wn, inv_wn, lane);
    auto prefetch = [&](int rr) {
        int x, y;
        orig_xy(rr, x, y);
        pf_cost = ccol[(size_t)rr * V + vl];
        pf_beta = scol[(size_t)rr * V + vl];
        pf_depth = dcol[rr];
        if (prp) {
            pf_prop = pcol[rr];
            pf_spread = qcol[rr];
        }
        {
            const f32x2 st2 = stcol[(size_t)rr * ststep];
            pf_stat = make_float2(st2.x, st2.y);
        }
        if constexpr (PF <= 8)
            load_col_window<PF>(cwin, rr, pf_win);
        else
            load_window<PF>(a, x, y, qx_col, qy_col, Wn, inv_wn, lane, pf_win);
    };
    if (Hs_rows > 1) prefetch(r);
    int buf = 0;
    float D = 0.0f;
    /* column windows of a fixed size fill their buffer without a branch */
    constexpr bool kWinAll = PF <= 8 && WN > 0 && kLanes * PF <= 2 * WN * WN;
    if (Hs_rows > 1) D = store_window<PF, kWinAll>(winbuf, pf_win, pf_stat.x, Wn, lane, need_d);

    unsigned long long n_evals = 0, n_samples = 0;
    if constexpr (kVec) {
        to_vgpr(n_evals);
        to_vgpr(n_samples);
    }
    int pb = 0; /* pass counter: parity selects the partial-sum buffer */

    /* one phase of hypotheses tview/tdep[0..nt) -> tout[0..nt); in the
     * small-view kernels the hypotheses' projections hypb[0..nt) */
    auto run_phase = [&](int nt, const float* win, float* tout, const PixelCtx& pc) {
        if constexpr (SM) {
            /* at most 3 x 7 hypotheses, each projection computed once, by one
             * lane, before the phase: the passes leave their window sums in
             * `part` (in the order a per-pass finish would combine them), and
             * one lane per hypothesis forms its 1 - NCC */
            for (int t0 = 0; t0 < nt; t0 += NGT) {
                const int ntp = min(NGT, nt - t0);
                const int lk = __builtin_ctz((unsigned)NGT) -
                               (ntp > 1 ? 32 - __builtin_clz((unsigned)(ntp - 1)) : 0);
                const Hyp q = load_hyp(hypb + 4 * (t0 + min((wave * NG + lane / G) >> lk, ntp - 1)));
                eval_pass<G, WN, NW, NB, true, true>(a, Hs, sb, win, wave, lane, ntp, nullptr,
                                                     nullptr, nullptr, part + t0 * NW, pc, &q);
            }
            lds_barrier();
            /* lane t of wave 0 finishes hypothesis t (t < nt <= 24); the
             * lanes past nt repeat the last one into entries nobody reads
             * (tout has 64), so there is no branch */
            if (wave == 0 && nt > 0) {
                const int t = min(lane, nt - 1);
                float4 s = part[t * NW];
                if constexpr (NW > 1) {
                    /* hypotheses split over the waves: partials in slice order */
                    const int ntp = min(NGT, nt - (t & ~(NGT - 1)));
                    const int lk = __builtin_ctz((unsigned)NGT) -
                                   (ntp > 1 ? 32 - __builtin_clz((unsigned)(ntp - 1)) : 0);
                    const int slices = (1 << lk) > NG ? (1 << lk) / NG : 1;
                    for (int w = 1; w < slices; w++) {
                        const float4 q = part[t * NW + w];
                        s.x += q.x;
                        s.y += q.y;
                        s.z += q.z;
                    }
                }
                tout[lane] = ncc_finish_sel(s.x, s.y, s.z, s.w, pc);
            }
            lds_barrier();
            return;
        }
        for (int t0 = 0; t0 < nt; t0 += NGT) {
            const int ntp = min(NGT, nt - t0);
            float4* pp = part + (pb & 1) * 64 * NW;
            const int kk = eval_pass<G, WN, NW, NB, true>(a, Hs, sb, win, wave, lane, ntp, tview + t0,
                                                tdep + t0, tout + t0, pp, pc);
            lds_barrier();
            if (kk > NG) combine_pass<NW>(lane, ntp, kk / NG, pp, tout + t0, pc);
            pb++;
        }
    };

    for (int k = 1; k < Hs_rows; k++, r += step) {
        float* win = winbuf + buf * WW;
        lds_barrier();
        /* This row's values, prefetched one row ago, are consumed before
         * the next row's loads go out: a first use after those loads would
         * wait for them as well (memory counters drain in order).  The
         * column-uniform ones move to scalar registers. */
        const float cst = pf_cost, beta = pf_beta;
        const float d_cur = uniform_f(pf_depth), spread = prp ? uniform_f(pf_spread) : 0.0f;
        float d_prop = prp ? uniform_f(pf_prop) : 0.0f;
        const float2 st = make_float2(uniform_f(pf_stat.x), uniform_f(pf_stat.y));

        int x, y;
        orig_xy(r, x, y);
        if (prp && d_prop > 0.0f && spread > 0.0f) d_prop = proposal_draw(d_prop, spread, x, y, a);
        PixelCtx pc;
        pixel_ctx(a, x, y, gl, pc);
        pc.mu = st.x;
        pc.c1 = st.y;
        pc.D = D;
        const int npx = pc.npx;
        /* The former depth equal to the current one: its NCC on a view is
         * the stored cost of that view (the same function of the same
         * pixel, depth and images), so former and current tie exactly and
         * findMinCost keeps the current depth unless the random (or
         * proposal) candidate is at least as good.  Only the left and top
         * border columns, where the reference's tile-index shift makes the
         * NCC depend on the sweep frame the stored cost came from, evaluate
         * the former depth anyway. */
        const bool same = d_former == d_cur && x >= a.hw && y >= a.hw;

        /* forward message with the stored costs, posterior, CDF
         * (src/patchMatch.cu:901-923) */
        const float g_prev = trans_g(m); /* shared with the refresh below */
        const float mf = g_prev + emission_lr(sentinel_emission(cst, is_view, lane, rng, vmask_v), mc);
        /* posterior log-odds of each view; the normalised CDF of its
         * probabilities (inclusive prefix sum over lanes 0..V-1 divided by
         * the total).  Where every posterior is below FP64's smallest
         * number (e^-744.4, only at tile-index-quirk costs) the reference's
         * arithmetic gives 0/0: the CDF is NaN and every draw takes the last
         * view.  Below e^-80 the probabilities are scaled by the largest. */
        float cdf;
        {
            const float L = mf + beta;
            float Lm = is_view ? L : -INFINITY;
            if (SM || V <= 16) {
                Lm = row_max16(Lm);
            } else {
#pragma unroll
                for (int o = 1; o < kLanes; o <<= 1) Lm = fmaxf(Lm, __shfl_xor(Lm, o, kLanes));
            }
            const float Lmax = lane_f(Lm, 0);
            float q = (Lmax > -80.0f) ? __builtin_amdgcn_rcpf(1.0f + __expf(-L)) : __expf(L - Lmax);
            q = is_view ? q : 0.0f;
            cdf = q;
            if (SM || V <= 16) {
                cdf = row_scan16f(cdf);
            } else {
#pragma unroll
                for (int o = 1; o < kLanes; o <<= 1) {
                    const float t = __shfl_up(cdf, o, kLanes);
                    if (lane >= o) cdf += t;
                }
            }
            const float total = lane_f(cdf, V - 1);
            cdf = (lane == V - 1) ? 1.0f : cdf * __builtin_amdgcn_rcpf(total);
            if (Lmax < -700.0f) {
                /* the reference's numerator m_f * beta of each view, in log */
                float U = is_view ? mf - __logf(1.0f + __expf(-beta)) : -INFINITY;
                if (SM || V <= 16) {
                    U = row_max16(U);
                } else {
#pragma unroll
                    for (int o = 1; o < kLanes; o <<= 1) U = fmaxf(U, __shfl_xor(U, o, kLanes));
                }
                if (lane_f(U, 0) < -744.4f) cdf = __builtin_nanf("");
            }
        }

        /* the next row's prefetch, once this row's values are in registers */
        __builtin_amdgcn_sched_barrier(0);
        if (k + 1 < Hs_rows) prefetch(r + step);

        /* draws: random depth, then S views (src/patchMatch.cu:937-961).
         * A view draw takes the first view whose CDF entry is >= u (none:
         * a NaN CDF, the last view). */
        float d_rand;
        uint64_t sel = 0;
        int cnt = 0;   /* draws that picked this lane's view */
        int my_id = 0; /* view of draw `lane` (draws 0..63) */
        if (SM || a.S < 8) {
            uint64_t ids6 = 0; /* the view ids, 6 bits per draw */
            /* The row's 1 + S draws unrolled, so the generator's word
             * rotation is register renaming.  u = k 2^-24 with the integer
             * k = (r >> 8) + 1, so u <= cdf is exactly k <= floor(cdf 2^24),
             * an integer compare whose lane mask is the draw's hits; lane
             * V - 1 holds 2^24 (its CDF entry is 1) unless the CDF is NaN,
             * where every entry converts to 0 and the draw takes view V - 1
             * through the bit set below. */
            const uint32_t k0 = (xw_next(rng) >> 8) + 1u;
            d_rand = (float)k0 * 5.9604644775390625e-08f * (a.far_d - a.near_d) + a.near_d;
            const uint32_t cvi = (is_view && cdf == cdf) ? (uint32_t)(cdf * 16777216.0f) : 0u;
            const uint64_t last = 1ull << (V - 1);
            auto draw = [&](int t) {
                const uint32_t kt = (xw_next(rng) >> 8) + 1u;
                /* the compare's lane mask as is (a ballot of the bool
                 * would select and compare it again) */
                const uint64_t hit = __builtin_amdgcn_uicmp(kt, cvi, 37 /* ICMP_ULE */) | last;
                const int id = (int)__builtin_ctzll(hit);
                sel |= 1ull << id;
                cnt += (lane == id) ? 1 : 0;
                ids6 |= (uint64_t)id << (6 * (t - 1));
            };
            if constexpr (SM) {
                /* a loop: no per-draw predicates held across the row loop;
                 * the ids are packed only when they are recorded */
                if (a.ids) {
                    for (int t = 1; t <= a.S; t++) draw(t);
                } else {
                    auto draw1 = [&]() {
                        const uint32_t kt = (xw_next(rng) >> 8) + 1u;
                        const uint64_t hit = __builtin_amdgcn_uicmp(kt, cvi, 37 /* ICMP_ULE */) | last;
                        const int id = (int)__builtin_ctzll(hit);
                        sel |= 1ull << id;
                        cnt += (lane == id) ? 1 : 0;
                    };
                    /* five draws at a time bring the generator's five words
                     * back to their registers (no moves) */
                    int t = 1;
                    for (; t + 4 <= a.S; t += 5) {
                        draw1();
                        draw1();
                        draw1();
                        draw1();
                        draw1();
                    }
                    for (; t <= a.S; t++) draw1();
                }
            } else {
#pragma unroll
                for (int t = 1; t < 8; t++)
                    if (t <= a.S) draw(t);
            }
            my_id = (int)((ids6 >> (6 * min(lane, 7))) & 63u);
        } else {
            /* floor(64 / V) draws are resolved by one ballot: lane d V + i
             * compares draw d with view i's CDF entry */
            d_rand = xw_uniform(rng) * (a.far_d - a.near_d) + a.near_d;
            const int per_b = kLanes / V;
            const int ld = lane / V, lv = lane - ld * V;
            const float cv = __shfl(cdf, lv, kLanes);
            const uint64_t vmask = (V >= 64) ? ~0ull : ((1ull << V) - 1ull);
            for (int j0 = 0; j0 < a.S; j0 += per_b) {
                const int nd = min(per_b, a.S - j0);
                float uj = 2.0f;
                for (int t = 0; t < nd; t++) {
                    const float u = xw_uniform(rng);
                    uj = (ld == t) ? u : uj;
                }
                const uint64_t hit = __ballot(ld < nd && uj <= cv);
                for (int t = 0; t < nd; t++) {
                    const uint64_t f = (hit >> (t * V)) & vmask;
                    const int id = f ? (int)__builtin_ctzll(f) : V - 1;
                    sel |= 1ull << id;
                    cnt += (lane == id) ? 1 : 0;
                    my_id = (lane == j0 + t) ? id : my_id;
                    if (j0 + t >= kLanes && lead && a.ids && lane == 0)
                        a.ids[((size_t)(j0 + t) * a.H + y) * a.ids_pitch + x] = (uint8_t)id;
                }
            }
        }
        /* the ids of draws 0..63, one per lane, in one store */
        if (lead && a.ids && lane < a.S)
            a.ids[((size_t)lane * a.H + y) * a.ids_pitch + x] = (uint8_t)my_id;
        const int U = __popcll(sel);
        const bool in_sel = is_view && ((sel >> lane) & 1ull);
        const int srank = __popcll(sel & ((1ull << lane) - 1ull));
        /* a learned proposal is a fourth candidate on the same views */
        const bool has_prop = prp && d_prop > 0.0f;
        const int per = (same ? 1 : 2) + (has_prop ? 1 : 0);
        /* Small-view kernels: lane 16 k + v projects view v at the former
         * (k = 0), random (1) or proposal (2) depth, every view and depth a
         * phase of this row can ask for; phase B's hypotheses are stored in
         * task order now, phase C's from the same registers once the winner
         * is known. */
        Hyp hq;
        const int hk = lane >> 4, hv = lane & 15;
        if constexpr (SM) {
            const float dk = hk == 0 ? d_former : (hk == 1 || !has_prop ? d_rand : d_prop);
            hq = make_hyp(Hs, sb, hv < V ? hv : 0, dk, pc);
            const bool on = lead && hv < V && ((sel >> hv) & 1ull) &&
                            (hk == 1 || (hk == 0 && !same) || (hk == 2 && has_prop));
            const int rk = __popc((uint32_t)sel & ((1u << hv) - 1u));
            const int pos = hk == 0 ? 0 : (same ? 0 : 1) + (hk == 2 ? 1 : 0);
            /* lanes without a hypothesis write four times to a float4 of
             * their own in the task table, which these kernels do not use
             * (no branch) */
            store_hyp(on ? hypb + 4 * (per * rk + pos) : (float4*)tview + lane, hq, on ? 1 : 0);
            lds_barrier();
        } else {
            if (in_sel) {
                int o = per * srank;
                if (!same) {
                    tview[o] = lane;
                    tdep[o++] = d_former;
                }
                tview[o] = lane;
                tdep[o++] = d_rand;
                if (has_prop) {
                    tview[o] = lane;
                    tdep[o] = d_prop;
                }
            }
            if (NW == 1) lds_barrier();
        }
        const int nt = per * U;

        /* phase B: former and random depth (and the proposal) on the sampled views */
        run_phase(nt, win, tres, pc);
        float nf = 0.0f, nr = 0.0f, np = 0.0f;
        if constexpr (SM) {
            /* every lane reads (per srank + 2 < 64, inside tres), the
             * selected views keep theirs */
            const int o = per * srank;
            const float t0 = tres[o], t1 = tres[o + 1], t2 = tres[o + 2];
            nf = in_sel ? (same ? cst : t0) : 0.0f;
            nr = in_sel ? (same ? t0 : t1) : 0.0f;
            np = in_sel && has_prop ? (same ? t1 : t2) : 0.0f;
        } else if (in_sel) {
            int o = per * srank;
            nf = same ? cst : tres[o++];
            nr = tres[o++];
            if (has_prop) np = tres[o];
        }
        /* candidate costs summed over the S draws: each view's cost times
         * the number of draws that picked it, summed over the views */
        float c0, c1s, c2, c3 = 0.0f;
        {
            const float k = is_view ? (float)cnt : 0.0f;
            auto vsum = [&](float v) { return SM ? lane_f(row_sum16(v), 0) : wave_sum(v, V); };
            c0 = vsum(is_view ? k * nf : 0.0f);
            c1s = vsum(is_view ? k * cst : 0.0f);
            c2 = vsum(is_view ? k * nr : 0.0f);
            if (has_prop) c3 = vsum(is_view ? k * np : 0.0f);
        }
        const float inv_s = 1.0f / (float)a.S;
        c0 *= inv_s;
        c1s *= inv_s;
        c2 *= inv_s;
        c3 *= inv_s;
        /* findMinCost (src/patchMatch.cu:613-618): strict <, ties to the
         * later candidate; the proposal comes last */
        int idx = c0 < c1s ? 0 : 1;
        float cidx = idx == 0 ? c0 : c1s;
        idx = cidx < c2 ? idx : 2;
        if (has_prop) {
            cidx = idx == 0 ? c0 : (idx == 1 ? c1s : c2);
            idx = cidx < c3 ? idx : 3;
        }
        const float best = idx == 0 ? d_former : (idx == 1 ? d_cur : (idx == 2 ? d_rand : d_prop));
        if (lead && lane == 0) dcol[r] = best;
        n_evals += (unsigned long long)nt;

        float cnew = cst;
        if (idx != 1) {
            /* phase C: remaining views at the winning depth */
            const uint64_t vmask = (V >= 64) ? ~0ull : ((1ull << V) - 1ull);
            const uint64_t rest = vmask & ~sel;
            const int nrest = V - U;
            const bool in_rest = is_view && !in_sel;
            const int rrank = __popcll(rest & ((1ull << lane) - 1ull));
            if constexpr (SM) {
                const int kc = idx == 0 ? 0 : (idx == 2 ? 1 : 2);
                const bool on = lead && hk == kc && hv < V && !((sel >> hv) & 1ull);
                const int rk = __popc((uint32_t)rest & ((1u << hv) - 1u));
                if constexpr (!kVec) hq = make_hyp(Hs, sb, hv < V ? hv : 0, best, pc);
                store_hyp(on ? hypb + 4 * rk : (float4*)tview + lane, hq, on ? 1 : 0);
                lds_barrier();
            } else {
                if (in_rest) {
                    tview[rrank] = lane;
                    tdep[rrank] = best;
                }
                if (NW == 1) lds_barrier();
            }
            run_phase(nrest, win, tres + kTasks, pc);
            if constexpr (SM) {
                const float tr = tres[kTasks + rrank]; /* rrank < 16 */
                cnew = in_sel ? ((idx == 0) ? nf : (idx == 2 ? nr : np)) : (in_rest ? tr : cnew);
            } else {
                if (in_sel) cnew = (idx == 0) ? nf : (idx == 2 ? nr : np);
                if (in_rest) cnew = tres[kTasks + rrank];
            }
            if (lead && is_view) ccol[(size_t)r * V + lane] = cnew;
            n_evals += (unsigned long long)nrest;
        }
        n_samples += (unsigned long long)npx * (unsigned long long)(nt + (idx != 1 ? V - U : 0));

        /* forward message with the updated costs (src/patchMatch.cu:996-999).
         * When the current depth won, the costs are unchanged: without a
         * sentinel (which draws again) the message is the one computed at
         * the top of the row, bit for bit. */
        if (idx == 1 && !sentinel_lanes(cst, vmask_v)) {
            m = mf;
        } else {
            m = g_prev + emission_lr(sentinel_emission(cnew, is_view, lane, rng, vmask_v), mc);
        }
        d_former = best;

        /* next window into the other buffer */
        if (k + 1 < Hs_rows) {
            buf ^= 1;
            D = store_window<PF, kWinAll>(winbuf + buf * WW, pf_win, pf_stat.x, Wn, lane, need_d);
        }
    }
    if (lead && lane == 0) {
        xw_store(rp, rng);
        if (a.counters) {
            atomicAdd(&a.counters[0], n_evals);
            atomicAdd(&a.counters[1], n_samples);
        }
    }
}

/* A 2048-column sweep puts 8 workgroups on each of the 256 CUs: with NW
 * waves each that is 2*NW waves per SIMD, all of which must be resident at
 * once (a column that waits for a free slot doubles the sweep time), so
 * the register budget is capped at 512 / WPE per lane, WPE = 2*NW.  A
 * 1536-column sweep with two waves per column needs 3 waves per SIMD
 * (WPE = 3), which leaves 168 registers instead of 128; its taps go out in
 * batches of 6 rows, which fit them without spilling (8-row batches spill
 * 6 registers and take 1.3% longer per 2048x1536 run). */
template <int G, int PF, int WN, int NW, bool H16, int WPE, bool SM = false, int ROT = -1,
          int PRP = -1>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE)))
void pmx_sweep_kernel(PmxSweepArgs a) {
    const int c = xcd_remap(blockIdx.x, gridDim.x);
    if (c >= a.Ws) return;
    sweep_column<G, PF, WN, NW, H16, (NW == 2 && WPE == 3) ? 6 : 0, SM, ROT, PRP>(a, c);
}

/*
 * One sweep step of many reference views in one persistent launch.  The
 * grid is sized to the workgroups the device keeps resident; each
 * workgroup takes column strips from the queue of its XCD until the queue
 * is empty.  Queue x holds, for every view in the order of tab (longest
 * columns first), the x-th of 8 contiguous ranges of that view's columns,
 * so neighbouring columns (which read neighbouring source footprints) stay
 * in one XCD's L2 and every XCD gets the same mix of column lengths.
 *   tab[v]            the sweep arguments of view v
 *   qstart[x*(n+1)+v] first queue slot of view v in queue x (qstart[.. n] = length)
 *   qcol[x*n+v]       first column of view v's range x
 *   queue[x]          next free slot of queue x (zero before the launch)
 * Workgroup b runs on XCD b % 8 (round-robin dispatch), which only decides
 * locality: every workgroup drains its queue, so every column is processed
 * whatever the placement.  WPE is the register budget in resident waves per
 * SIMD (512 / WPE registers per lane): 2 * NW keeps every register the
 * column code wants; 3 with one wave per strip puts a third column in
 * flight on each SIMD, which a batch (with far more columns than slots)
 * can fill.  In 168 registers a whole window column of taps in flight
 * spills, so those taps go out in batches of NB rows: 8 with FP16 texels,
 * 5 with FP32 (the fastest on the 16-view 10-source workload; 4 to 16
 * measured).
 */
template <int G, int PF, int WN, int NW, bool H16, int WPE, bool SM = false, int ROT = -1,
          int PRP = -1>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE)))
void pmx_sweep_batch_kernel(const PmxSweepArgs* __restrict__ tab, const int* __restrict__ qstart,
                            const int* __restrict__ qcol, int n, unsigned* __restrict__ queue) {
    constexpr int NB = (WPE == 3 && NW == 1) ? (H16 ? 8 : 5) : 0;
    __shared__ int s_slot;
    const int x = (int)(blockIdx.x & 7u);
    const int* __restrict__ st = qstart + x * (n + 1);
    const int total = st[n];
    for (;;) {
        if (threadIdx.x == 0) s_slot = (int)atomicAdd(queue + x, 1u);
        __syncthreads();
        const int j = s_slot;
        __syncthreads();
        if (j >= total) break;
        int v = 0;
        while (st[v + 1] <= j) v++;
        sweep_column<G, PF, WN, NW, H16, NB, SM, ROT, PRP>(tab[v], qcol[x * n + v] + (j - st[v]));
        __syncthreads();
    }
}

inline dim3 grid2(int w, int h, dim3 b) { return dim3((w + b.x - 1) / b.x, (h + b.y - 1) / b.y); }

} // namespace

/* ----------------------------------------------------------- launchers */

hipError_t pmx_launch_transpose_f32(hipStream_t s, const float* in, float* out, int width,
                                    int height, int ld_in, int ld_out) {
    dim3 b(32, 8), g((width + 31) / 32, (height + 31) / 32);
    hipLaunchKernelGGL(transpose_f32_kernel, g, b, 0, s, in, out, width, height, ld_in, ld_out);
    return hipGetLastError();
}

hipError_t pmx_launch_transpose_rec(hipStream_t s, const float* in, float* out, int width,
                                    int height, int E, int ld_in, int ld_out) {
    if (E == 1) return pmx_launch_transpose_f32(s, in, out, width, height, ld_in, ld_out);
    const int T = E > 32 ? 8 : 16;
    if (E < 1 || T * E > 512) return hipErrorInvalidValue; /* one tile row per <= 2 floats a thread */
    dim3 b(256), g((width + T - 1) / T, (height + T - 1) / T);
    const size_t sh = sizeof(float) * T * (T * E + 1);
    hipLaunchKernelGGL(transpose_rec_kernel, g, b, sh, s, in, out, width, height, E, T, ld_in,
                       ld_out);
    return hipGetLastError();
}

hipError_t pmx_launch_rng_init(hipStream_t s, pmx_rng_state* st, int pitch, int width, int height,
                               int bx, int by, uint32_t seed) {
    dim3 b(32, 8);
    hipLaunchKernelGGL(rng_init_kernel, grid2(width, height, b), b, 0, s, st, pitch, width, height,
                       bx, by, seed);
    return hipGetLastError();
}

hipError_t pmx_launch_fill_depth(hipStream_t s, pmx_rng_state* st, int st_pitch, float* out,
                                 int out_ld, int width, int height, float near_d, float far_d,
                                 int mode) {
    dim3 b(32, 8);
    hipLaunchKernelGGL(fill_depth_kernel, grid2(width, height, b), b, 0, s, st, st_pitch, out,
                       out_ld, width, height, near_d, far_d, mode);
    return hipGetLastError();
}

hipError_t pmx_launch_blur(hipStream_t s, float* data, int ld, int width, int height, int depth,
                           long layer_stride, const float* taps, int ntaps, float* tmp) {
    Taps t{};
    for (int i = 0; i < ntaps && i < 45; i++) t.t[i] = taps[i];
    const int fr = ntaps >> 1;
    dim3 b(32, 8), g = grid2(width, height, b);
    for (int k = 0; k < depth; k++) {
        float* layer = data + (size_t)k * layer_stride;
        hipLaunchKernelGGL(blur_rows_kernel, g, b, 0, s, layer, tmp, ld, width, height, t, fr);
        hipLaunchKernelGGL(blur_cols_kernel, g, b, 0, s, tmp, layer, ld, width, height, t, fr);
    }
    return hipGetLastError();
}

hipError_t pmx_launch_box_sums(hipStream_t s, const float* img, int img_ld, int width, int height,
                               int hw, float* out_I, float* out_sI, float* out_sII, int out_ld,
                               float2* stats, int stats_ld, double* tmp) {
    dim3 b(32, 8), g = grid2(width, height, b);
    hipLaunchKernelGGL(box_rows_kernel, g, b, 0, s, img, img_ld, width, height, hw, (double2*)tmp);
    hipLaunchKernelGGL(box_cols_kernel, g, b, 0, s, (const double2*)tmp, img, img_ld, width, height,
                       hw, out_I, out_sI, out_sII, out_ld, stats, stats_ld);
    return hipGetLastError();
}

hipError_t pmx_launch_compose(hipStream_t s, const double* cams, int V, double oX, double oZ,
                              double* H_out) {
    hipLaunchKernelGGL(compose_kernel, dim3((V + 63) / 64), dim3(64), 0, s, cams, V, oX, oZ, H_out);
    return hipGetLastError();
}

hipError_t pmx_launch_convert_u8(hipStream_t s, const uint8_t* in, int in_pitch, float* out,
                                 int out_ld, int width, int height) {
    dim3 b(32, 8);
    hipLaunchKernelGGL(convert_u8_kernel, grid2(width, height, b), b, 0, s, in, in_pitch, out,
                       out_ld, width, height);
    return hipGetLastError();
}

hipError_t pmx_launch_gather_costs(hipStream_t s, const float* cost, int cost_ld, float* out,
                                   int out_ld, int width, int height, int V) {
    dim3 b(32, 8);
    hipLaunchKernelGGL(gather_costs_kernel, grid2(width, height, b), b, 0, s, cost, cost_ld, out,
                       out_ld, width, height, V);
    return hipGetLastError();
}

hipError_t pmx_launch_scatter_costs(hipStream_t s, const float* in, int in_ld, float* cost,
                                    int cost_ld, int width, int height, int V) {
    dim3 b(32, 8);
    hipLaunchKernelGGL(scatter_costs_kernel, grid2(width, height, b), b, 0, s, in, in_ld, cost,
                       cost_ld, width, height, V);
    return hipGetLastError();
}

namespace {
template <int G, int PF, int WN>
hipError_t launch_cost_all_t(hipStream_t s, const PmxSweepArgs& a) {
    const int Wn = 2 * a.hw + 1;
    const size_t sh = sizeof(double) * a.V * 18 + sizeof(float4) * (128 + 4 * a.V) +
                      sizeof(float) * (2 * Wn * Wn);
    const long tiles = (long)((a.W + kCostTW - 1) / kCostTW) * ((a.H + kCostTH - 1) / kCostTH);
    const dim3 grid((unsigned)(tiles * kCostTW * kCostTH));
    if (a.src_h16)
        hipLaunchKernelGGL((pmx_cost_all_kernel<G, PF, WN, true>), grid, dim3(64), sh, s, a);
    else
        hipLaunchKernelGGL((pmx_cost_all_kernel<G, PF, WN, false>), grid, dim3(64), sh, s, a);
    return hipGetLastError();
}

template <int G, int PF, int WN>
hipError_t launch_volume_t(hipStream_t s, const PmxSweepArgs& a, int v0, int v1, uint16_t* out,
                           int W0, int H0) {
    const int Wn = 2 * a.hw + 1;
    const size_t sh = sizeof(double) * a.V * 18 + sizeof(float4) * (128 + 4 * 64) +
                      sizeof(float) * (64 + 2 * Wn * Wn);
    const dim3 grid((unsigned)((long)W0 * H0));
    if (a.src_h16)
        hipLaunchKernelGGL((pmx_volume_kernel<G, PF, WN, true>), grid, dim3(64), sh, s, a, v0, v1,
                           out, W0, H0);
    else
        hipLaunchKernelGGL((pmx_volume_kernel<G, PF, WN, false>), grid, dim3(64), sh, s, a, v0, v1,
                           out, W0, H0);
    return hipGetLastError();
}

template <int G, int PF, int WN, int NW, int WPE = (2 * NW > 8 ? 8 : 2 * NW), bool SM = false,
          int ROT = -1, int PRP = -1>
hipError_t launch_sweep_t(hipStream_t s, const PmxSweepArgs& a) {
    const int Wn = 2 * a.hw + 1;
    const size_t sh = sizeof(double) * a.V * 18 + sizeof(float4) * (2 * 64 * NW + (SM ? 4 * kHypMax : 0)) +
                      sizeof(float) * (4 * Wn * Wn + 3 * kTasks + 64 + 64);
    if (a.src_h16)
        hipLaunchKernelGGL((pmx_sweep_kernel<G, PF, WN, NW, true, WPE, SM, ROT, PRP>), dim3(a.Ws),
                           dim3(64 * NW), sh, s, a);
    else
        hipLaunchKernelGGL((pmx_sweep_kernel<G, PF, WN, NW, false, WPE, SM, ROT, PRP>), dim3(a.Ws),
                           dim3(64 * NW), sh, s, a);
    return hipGetLastError();
}

/* The specialisation for at most 16 views and 7 draws per row, with the
 * sweep frame and the proposal's presence fixed at compile time. */
template <int G, int PF, int WN, int NW, int WPE>
hipError_t launch_sweep_sm(hipStream_t s, const PmxSweepArgs& a) {
    const bool p = a.prop != nullptr;
    if (a.rotated)
        return p ? launch_sweep_t<G, PF, WN, NW, WPE, true, 1, 1>(s, a)
                 : launch_sweep_t<G, PF, WN, NW, WPE, true, 1, 0>(s, a);
    return p ? launch_sweep_t<G, PF, WN, NW, WPE, true, 0, 1>(s, a)
             : launch_sweep_t<G, PF, WN, NW, WPE, true, 0, 0>(s, a);
}
template <int G, int PF, int WN, int NW, int WPE = (2 * NW > 8 ? 8 : 2 * NW), bool SM = false,
          int ROT = -1, int PRP = -1>
hipError_t launch_sweep_batch_t(hipStream_t s, const PmxBatchLaunch& b) {
    const int Wn = 2 * b.hw + 1;
    const size_t sh = sizeof(double) * b.max_V * 18 + sizeof(float4) * (2 * 64 * NW + (SM ? 4 * kHypMax : 0)) +
                      sizeof(float) * (4 * Wn * Wn + 3 * kTasks + 64 + 64);
    if (b.h16)
        hipLaunchKernelGGL((pmx_sweep_batch_kernel<G, PF, WN, NW, true, WPE, SM, ROT, PRP>), dim3(b.grid),
                           dim3(64 * NW), sh, s, b.tab, b.qstart, b.qcol, b.n, b.queue);
    else
        hipLaunchKernelGGL((pmx_sweep_batch_kernel<G, PF, WN, NW, false, WPE, SM, ROT, PRP>), dim3(b.grid),
                           dim3(64 * NW), sh, s, b.tab, b.qstart, b.qcol, b.n, b.queue);
    return hipGetLastError();
}

template <int G, int PF, int WN, int NW, int WPE = (2 * NW > 8 ? 8 : 2 * NW), bool SM = false,
          int ROT = -1, int PRP = -1>
hipError_t occupancy_batch_t(const PmxBatchLaunch& b, int* per_cu) {
    const int Wn = 2 * b.hw + 1;
    const size_t sh = sizeof(double) * b.max_V * 18 + sizeof(float4) * (2 * 64 * NW + (SM ? 4 * kHypMax : 0)) +
                      sizeof(float) * (4 * Wn * Wn + 3 * kTasks + 64 + 64);
    if (b.h16)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            per_cu, pmx_sweep_batch_kernel<G, PF, WN, NW, true, WPE, SM, ROT, PRP>, 64 * NW, sh);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        per_cu, pmx_sweep_batch_kernel<G, PF, WN, NW, false, WPE, SM, ROT, PRP>, 64 * NW, sh);
}

/* Batched sweeps of views with at most 16 sources and 7 draws per row: the
 * sweep frame and the proposal's presence fixed at compile time. */
template <int G, int PF, int WN, int NW, int WPE>
hipError_t batch_sm(hipStream_t s, const PmxBatchLaunch& b, int* per_cu) {
#define PMX_BATCH_SM(R, P)                                                              \
    return per_cu ? occupancy_batch_t<G, PF, WN, NW, WPE, true, R, P>(b, per_cu)       \
                  : launch_sweep_batch_t<G, PF, WN, NW, WPE, true, R, P>(s, b);
    if (b.rot) {
        if (b.prp) { PMX_BATCH_SM(1, 1) }
        PMX_BATCH_SM(1, 0)
    }
    if (b.prp) { PMX_BATCH_SM(0, 1) }
    PMX_BATCH_SM(0, 0)
#undef PMX_BATCH_SM
}
} // namespace

/* Batched sweeps: the same shapes as single-view sweeps with one or two
 * waves per column strip; shape 3 is one wave per strip at three resident
 * waves per SIMD (the configured window sizes only; others run shape 1).
 * per_cu != NULL: report the resident workgroups per CU instead of
 * launching. */
hipError_t pmx_launch_sweep_batch(hipStream_t s, const PmxBatchLaunch& b, int* per_cu) {
    const int Wn = 2 * b.hw + 1;
    if (b.sm) {
        if (b.waves == 3 && Wn == 15) return batch_sm<16, 4, 15, 1, 3>(s, b, per_cu);
        if (b.waves == 3 && Wn == 11) return batch_sm<16, 2, 11, 1, 3>(s, b, per_cu);
        if (b.waves == 1 && Wn == 15) return batch_sm<16, 4, 15, 1, 2>(s, b, per_cu);
        if (b.waves == 1 && Wn == 11) return batch_sm<16, 2, 11, 1, 2>(s, b, per_cu);
    }
    if (b.waves == 3 && Wn == 15)
        return per_cu ? occupancy_batch_t<16, 4, 15, 1, 3>(b, per_cu)
                      : launch_sweep_batch_t<16, 4, 15, 1, 3>(s, b);
    if (b.waves == 3 && Wn == 11)
        return per_cu ? occupancy_batch_t<16, 2, 11, 1, 3>(b, per_cu)
                      : launch_sweep_batch_t<16, 2, 11, 1, 3>(s, b);
#define PMX_BATCH_NW(G, PF, WN)                                                              \
    if (b.waves == 2)                                                                         \
        return per_cu ? occupancy_batch_t<G, PF, WN, 2>(b, per_cu)                           \
                      : launch_sweep_batch_t<G, PF, WN, 2>(s, b);                            \
    return per_cu ? occupancy_batch_t<G, PF, WN, 1>(b, per_cu)                               \
                  : launch_sweep_batch_t<G, PF, WN, 1>(s, b);
    if (Wn == 15) { PMX_BATCH_NW(16, 4, 15) }
    if (Wn == 11) { PMX_BATCH_NW(16, 2, 11) }
    if (Wn <= 16) { PMX_BATCH_NW(16, 4, 0) }
    if (Wn <= 32) { PMX_BATCH_NW(32, 16, 0) }
    PMX_BATCH_NW(64, 44, 0)
#undef PMX_BATCH_NW
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0u;
}

hipError_t pmx_launch_zero_u32(hipStream_t s, unsigned* p, int n) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

/* Window sizes 11 and 15 (the configured workloads) get fully unrolled
 * row loops; every other size runs the same code with a dynamic trip
 * count. */
hipError_t pmx_launch_cost_all(hipStream_t s, const PmxSweepArgs& a) {
    const int Wn = 2 * a.hw + 1;
    if (Wn == 15) return launch_cost_all_t<16, 4, 15>(s, a);
    if (Wn == 11) return launch_cost_all_t<16, 2, 11>(s, a);
    if (Wn <= 16) return launch_cost_all_t<16, 4, 0>(s, a);
    if (Wn <= 32) return launch_cost_all_t<32, 16, 0>(s, a);
    return launch_cost_all_t<64, 44, 0>(s, a);
}

hipError_t pmx_launch_cost_volume(hipStream_t s, const PmxSweepArgs& a, int view0, int view1,
                                  uint16_t* out, int W0, int H0) {
    const int Wn = 2 * a.hw + 1;
    if (Wn == 15) return launch_volume_t<16, 4, 15>(s, a, view0, view1, out, W0, H0);
    if (Wn == 11) return launch_volume_t<16, 2, 11>(s, a, view0, view1, out, W0, H0);
    if (Wn <= 16) return launch_volume_t<16, 4, 0>(s, a, view0, view1, out, W0, H0);
    if (Wn <= 32) return launch_volume_t<32, 16, 0>(s, a, view0, view1, out, W0, H0);
    return launch_volume_t<64, 44, 0>(s, a, view0, view1, out, W0, H0);
}

hipError_t pmx_launch_pair_f32(hipStream_t s, const float* data, float2* out, size_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pair_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, data,
                       out, n);
    return hipGetLastError();
}

hipError_t pmx_launch_quantize_f16(hipStream_t s, float* data, uint16_t* out, size_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(quantize_f16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, data,
                       out, n);
    return hipGetLastError();
}

hipError_t pmx_launch_seed_depth(hipStream_t s, const float* prop, int prop_ld, float* depth,
                                 int depth_ld, int width, int height) {
    dim3 b(32, 8);
    hipLaunchKernelGGL(seed_depth_kernel, grid2(width, height, b), b, 0, s, prop, prop_ld, depth,
                       depth_ld, width, height);
    return hipGetLastError();
}

hipError_t pmx_launch_sweep(hipStream_t s, const PmxSweepArgs& a, int waves) {
    const int Wn = 2 * a.hw + 1;
    /* at most 16 views and 7 draws per row (the configured workloads): the
     * specialisation without the wide-view and many-draw paths */
    if (a.V <= 16 && a.S <= 7) {
        if (waves == 1 && Wn == 15) return launch_sweep_sm<16, 4, 15, 1, 2>(s, a);
        if (waves == 1 && Wn == 11) return launch_sweep_sm<16, 2, 11, 1, 2>(s, a);
        if (waves == 3 && Wn == 15) return launch_sweep_sm<16, 4, 15, 2, 3>(s, a);
        if (waves == 3 && Wn == 11) return launch_sweep_sm<16, 2, 11, 2, 3>(s, a);
    }
#define PMX_SWEEP_NW(G, PF, WN)                                      \
    switch (waves) {                                                 \
    case 1: return launch_sweep_t<G, PF, WN, 1>(s, a);               \
    case 4: return launch_sweep_t<G, PF, WN, 4>(s, a);               \
    default: return launch_sweep_t<G, PF, WN, 2>(s, a);              \
    }
    /* waves == 3: two waves per column at three waves per SIMD (the
     * configured window sizes only) */
    if (waves == 3 && Wn == 15) return launch_sweep_t<16, 4, 15, 2, 3>(s, a);
    if (waves == 3 && Wn == 11) return launch_sweep_t<16, 2, 11, 2, 3>(s, a);
    if (waves == 3) waves = 1;
    if (Wn == 15) { PMX_SWEEP_NW(16, 4, 15) }
    if (Wn == 11) { PMX_SWEEP_NW(16, 2, 11) }
    if (Wn <= 16) { PMX_SWEEP_NW(16, 4, 0) }
    if (Wn <= 32) { PMX_SWEEP_NW(32, 16, 0) }
    PMX_SWEEP_NW(64, 44, 0)
#undef PMX_SWEEP_NW
}
