/*
 * pmx_oracle.cpp -- scalar FP64 CPU restatement of the reference PatchMatch
 * kernels (yhsa678/patchMatch_cuda).  TEST INFRASTRUCTURE ONLY: it is the
 * checker for the HIP engine and the CPU baseline of the benchmark; no
 * product code links or calls it.
 *
 * Columns of a sweep are independent (each owns its generator state), so
 * the sweeps run with an OpenMP parallel-for over columns and produce the
 * same result for any thread count.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   rng          - golden vectors (tests/golden/rng_vectors.json) checked
 *                  against an independent numpy restatement of xorwow and
 *                  SplitMix64; block-id layout against hand-computed ids.
 *   init depth   - closed form of the uniform map on the pinned draws.
 *   blur taps    - closed forms (Gaussian exp(-i^2/2s^2)/sum, box 1/width);
 *                  impulse response recovers the taps; constants are fixed.
 *   box sums     - brute-force window sums; constant image gives c*count.
 *   transpose    - permutation check on odd sizes.
 *   homography   - 3D points on the plane project to the same source pixel
 *                  through H and through the two cameras; identity cameras.
 *   NCC          - identical patch -> 0, affine a>0 -> 0, a<0 -> 2,
 *                  constant source -> 1, out of frame -> 2, brute-force
 *                  restatement on random windows incl. the left-border
 *                  tile-index quirk (Q1), rotated frame == transposed data.
 *   messages     - two-state forward/backward closed forms and brute-force
 *                  enumeration of state paths for the posterior.
 *   view draw    - CDF inverse on hand-computed cases incl. u == cdf.
 *   sweep / run  - converged depth against ground truth on a synthetic
 *                  scene; identical output for 1 and N threads.
 */
#include "pmx_oracle.h"

#include <atomic>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Constants the reference writes as FP32 literals, at their FP32 values.
 * src/patchMatch.cu:9-10 (FIX_STATE_PROB, CHANGE_STATE_PROB), :673 (x). */
const double kFix = (double)0.999f;
const double kChange = (double)(1.0f - 0.999f);
const double kUniformX = (double)0.83f;
const double kSentinel = 2.0;     /* src/patchMatch.cu:594-595, :680 */
const double kVarFloor = 0.00001; /* src/patchMatch.cu:599-600 */

int g_threads = 0; /* 0 = OpenMP default */

inline uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline int nthreads() {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

void mat3_mul(const double* a, const double* b, double* out) {
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += a[i * 3 + k] * b[k * 3 + j];
            t[i * 3 + j] = s;
        }
    std::memcpy(out, t, sizeof(t));
}

void mat3_inv(const double* m, double* out) {
    const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5],
                 g = m[6], h = m[7], i = m[8];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    const double r = 1.0 / det;
    out[0] = A * r;
    out[1] = -(b * i - c * h) * r;
    out[2] = (b * f - c * e) * r;
    out[3] = B * r;
    out[4] = (a * i - c * g) * r;
    out[5] = -(a * f - c * d) * r;
    out[6] = C * r;
    out[7] = -(a * h - b * g) * r;
    out[8] = (a * e - b * d) * r;
}

} // namespace

struct pmxo_problem {
    pmxo_params p;
    std::vector<double> ref;    /* W*H */
    std::vector<double> srcs;   /* V*srcH*srcW */
    std::vector<double> H;      /* V*18 */
    std::vector<double> sum_I;  /* W*H */
    std::vector<double> sum_II; /* W*H */
    /* learned proposal (depth, spread), original frame (W*H) and rotated
     * frame (H*W); empty = none */
    std::vector<double> prop, prop_t, spread, spread_t;
    mutable int sweep_id = 0; /* sweeps since the proposal was set */
    int prop_sweeps = 0;      /* sweeps that get the proposal (0: all) */
    mutable std::atomic<uint64_t> ncc_evals{0};
    mutable std::atomic<uint64_t> samples{0};
};

extern "C" {

/* ------------------------------------------------------------------ rng */

void pmxo_rng_seed(pmxo_rng* st, uint32_t id, uint32_t seed) {
    uint64_t s = ((uint64_t)seed << 32) | (uint64_t)id;
    const uint64_t a = splitmix64(&s), b = splitmix64(&s), c = splitmix64(&s);
    st->x = (uint32_t)a;
    st->y = (uint32_t)(a >> 32);
    st->z = (uint32_t)b;
    st->w = (uint32_t)(b >> 32);
    st->v = (uint32_t)c;
    st->d = (uint32_t)(c >> 32);
    if ((st->x | st->y | st->z | st->w | st->v) == 0u) st->x = 1u;
}

uint32_t pmxo_rng_next(pmxo_rng* st) {
    const uint32_t t = st->x ^ (st->x >> 2);
    st->x = st->y;
    st->y = st->z;
    st->z = st->w;
    st->w = st->v;
    st->v = (st->v ^ (st->v << 4)) ^ (t ^ (t << 1));
    st->d += 362437u;
    return st->d + st->v;
}

double pmxo_rng_uniform(pmxo_rng* st) {
    const uint32_t r = pmxo_rng_next(st);
    return (double)((r >> 8) + 1u) * (1.0 / 16777216.0);
}

/* src/array2d_psng.cu:8-9: id = (by*gx + bx)*bdx*bdy + ty*bdx + tx. */
uint32_t pmxo_seed_id(int x, int y, int width, int block_x, int block_y) {
    const int gx = (width - 1) / block_x + 1;
    const int bx = x / block_x, tx = x % block_x;
    const int by = y / block_y, ty = y % block_y;
    return (uint32_t)((by * gx + bx) * block_x * block_y + ty * block_x + tx);
}

void pmxo_rng_init_grid(pmxo_rng* states, int width, int height, int block_x,
                        int block_y, uint32_t seed) {
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++)
            pmxo_rng_seed(&states[(size_t)y * width + x],
                          pmxo_seed_id(x, y, width, block_x, block_y), seed);
}

/* src/array2d_wrapper.cu:15: (1 - u)(far - near) + near. */
void pmxo_init_depth(pmxo_rng* states, int width, int height, double near_d,
                     double far_d, int mode, double* depth) {
    for (size_t i = 0; i < (size_t)width * height; i++) {
        const double u = pmxo_rng_uniform(&states[i]);
        if (mode == 1) {
            const double lo = 1.0 / far_d, hi = 1.0 / near_d;
            depth[i] = 1.0 / ((1.0 - u) * (hi - lo) + lo);
        } else {
            depth[i] = (1.0 - u) * (far_d - near_d) + near_d;
        }
    }
}

/* -------------------------------------------------------- image helpers */

/* src/GaussianBlurCUDA.cu:99-130: width = int(5*sigma) made odd, clamped to
 * [5, 45]. */
int pmxo_blur_taps(double sigma, int mode, double* taps) {
    int width = (int)(5.0f * (float)sigma);
    if (width % 2 == 0) width += 1;
    int sz = (width - 1) >> 1;
    if (width > 45) {
        sz = 45 >> 1;
        width = 45;
    } else if (width < 5) {
        sz = 5 >> 1;
        width = 5;
    }
    if (mode == 2) {
        for (int i = 0; i < width; i++) taps[i] = 1.0 / (double)width;
        return width;
    }
    const double rv = -0.5 / (sigma * sigma);
    double ksum = 0.0;
    for (int i = -sz; i <= sz; i++) {
        taps[i + sz] = std::exp((double)(i * i) * rv);
        ksum += taps[i + sz];
    }
    for (int i = 0; i < width; i++) taps[i] /= ksum;
    return width;
}

/* src/GaussianBlurCUDA.cu:39-70 (row pass then column pass, clamp-to-edge
 * addressing of the texture reference defaults). */
void pmxo_blur(const double* in, int width, int height, double sigma, int mode,
               double* out) {
    if (mode == 0) { /* no pre-blur: the reference's own behaviour (the filter is never called) */
        std::copy(in, in + (size_t)width * height, out);
        return;
    }
    double taps[45];
    const int tw = pmxo_blur_taps(sigma, mode, taps);
    const int fr = tw >> 1;
    std::vector<double> tmp((size_t)width * height);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            double s = 0.0;
            for (int k = -fr; k <= fr; k++) {
                int xx = x + k;
                xx = xx < 0 ? 0 : (xx > width - 1 ? width - 1 : xx);
                s += in[(size_t)y * width + xx] * taps[fr - k];
            }
            tmp[(size_t)y * width + x] = s;
        }
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            double s = 0.0;
            for (int k = -fr; k <= fr; k++) {
                int yy = y + k;
                yy = yy < 0 ? 0 : (yy > height - 1 ? height - 1 : yy);
                s += tmp[(size_t)yy * width + x] * taps[fr - k];
            }
            out[(size_t)y * width + x] = s;
        }
}

/* src/array2d_refImg.cu:11-67: horizontal then vertical sums, zero border. */
void pmxo_box_sums(const double* img, int width, int height, int hw,
                   double* sum_I, double* sum_II) {
    std::vector<double> r1((size_t)width * height), r2((size_t)width * height);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            double a = 0.0, b = 0.0;
            for (int k = -hw; k <= hw; k++) {
                const int xx = x + k;
                if (xx < 0 || xx >= width) continue;
                const double v = img[(size_t)y * width + xx];
                a += v;
                b += v * v;
            }
            r1[(size_t)y * width + x] = a;
            r2[(size_t)y * width + x] = b;
        }
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            double a = 0.0, b = 0.0;
            for (int k = -hw; k <= hw; k++) {
                const int yy = y + k;
                if (yy < 0 || yy >= height) continue;
                a += r1[(size_t)yy * width + x];
                b += r2[(size_t)yy * width + x];
            }
            sum_I[(size_t)y * width + x] = a;
            sum_II[(size_t)y * width + x] = b;
        }
}

/* src/cudaTranspose.h:8-46: a pure permutation. */
void pmxo_transpose(const double* in, int width, int height, double* out) {
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++)
            out[(size_t)x * height + y] = in[(size_t)y * width + x];
}

/* src/image.cpp:4-13. */
void pmxo_compose(const double* Kr, const double* Rr, const double* Cr,
                  const double* Ks, const double* Rs, const double* Cs,
                  double oX, double oZ, double* H1, double* H2) {
    double RrT[9], Rrel[9], Kri[9], t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) RrT[i * 3 + j] = Rr[j * 3 + i];
    mat3_mul(Rs, RrT, Rrel);
    double dc[3] = {Cs[0] - Cr[0], Cs[1] - Cr[1], Cs[2] - Cr[2]};
    double trel[3];
    for (int i = 0; i < 3; i++)
        trel[i] = Rs[i * 3 + 0] * dc[0] + Rs[i * 3 + 1] * dc[1] + Rs[i * 3 + 2] * dc[2];
    mat3_inv(Kr, Kri);
    mat3_mul(Ks, Rrel, t);
    mat3_mul(t, Kri, H1);
    const double n[3] = {oX, 0.0, oZ};
    double tn[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) tn[i * 3 + j] = trel[i] * n[j];
    mat3_mul(Ks, tn, t);
    mat3_mul(t, Kri, H2);
}

/* ----------------------------------------------------------- problem */

pmxo_problem* pmxo_problem_create(const pmxo_params* p, const double* ref,
                                  const double* srcs, const double* H1H2) {
    pmxo_problem* pb = new pmxo_problem();
    pb->p = *p;
    const size_t npx = (size_t)p->W * p->H;
    pb->ref.assign(ref, ref + npx);
    pb->srcs.assign(srcs, srcs + (size_t)p->V * p->srcW * p->srcH);
    pb->H.assign(H1H2, H1H2 + (size_t)p->V * 18);
    pb->sum_I.resize(npx);
    pb->sum_II.resize(npx);
    pmxo_box_sums(pb->ref.data(), p->W, p->H, p->hw, pb->sum_I.data(), pb->sum_II.data());
    return pb;
}

void pmxo_problem_destroy(pmxo_problem* pb) { delete pb; }

void pmxo_problem_ref_sums(const pmxo_problem* pb, double* sum_I, double* sum_II) {
    std::memcpy(sum_I, pb->sum_I.data(), pb->sum_I.size() * sizeof(double));
    std::memcpy(sum_II, pb->sum_II.data(), pb->sum_II.size() * sizeof(double));
}

} // extern "C"

namespace {

/* Layered texture fetch with linear filtering, border addressing and
 * unnormalised coordinates (src/patchMatch.cu:192-195): the sample at
 * texel-space position (X, Y) with texel centres on integers. */
inline double src_bilinear(const pmxo_problem* pb, int view, double X, double Y) {
    const int sw = pb->p.srcW, sh = pb->p.srcH;
    if (!(X > -1.0 && X < (double)sw && Y > -1.0 && Y < (double)sh)) return 0.0;
    const double fx = std::floor(X), fy = std::floor(Y);
    const int ix = (int)fx, iy = (int)fy;
    const double a = X - fx, b = Y - fy;
    const double* img = pb->srcs.data() + (size_t)view * sw * sh;
    auto tex = [&](int i, int j) -> double {
        return (i >= 0 && i < sw && j >= 0 && j < sh) ? img[(size_t)j * sw + i] : 0.0;
    };
    return (1.0 - a) * (1.0 - b) * tex(ix, iy) + a * (1.0 - b) * tex(ix + 1, iy) +
           (1.0 - a) * b * tex(ix, iy + 1) + a * b * tex(ix + 1, iy + 1);
}

/* Reference value in the sweep frame (refImgTexture at texel centres,
 * border 0; rotated frames swap x and y, src/patchMatch.cu:634-659). */
inline double ref_s(const pmxo_problem* pb, int row_s, int col_s, int rotated) {
    const int W = pb->p.W, H = pb->p.H;
    const int x = rotated ? row_s : col_s;
    const int y = rotated ? col_s : row_s;
    if (x < 0 || x >= W || y < 0 || y >= H) return 0.0;
    return pb->ref[(size_t)y * W + x];
}

/* scale = n . K^-1 (x + 0.5, y + 0.5, 1) with the reference's four K^-1
 * entries and orientation[1] == 0 (src/patchMatch.cu:798-800, :876-883). */
inline double plane_scale(const pmxo_problem* pb, int row_s, int col_s, int rotated) {
    const pmxo_params& p = pb->p;
    const double x = rotated ? row_s : col_s;
    const double y = rotated ? col_s : row_s;
    return p.oX * (p.invK[0] * (x + 0.5) + p.invK[1]) +
           0.0 * (p.invK[2] * (y + 0.5) + p.invK[3]) + p.oZ;
}

inline bool near_floor(double v) { return std::fabs(v - kVarFloor) < 1e-3 * kVarFloor; }

/* computeNCC (src/patchMatch.cu:433-611) plus readImageIntoSharedMemory
 * (:620-667): the reference tile index advances only on in-bounds window
 * columns (:536-551), which this restatement keeps (counter j). */
double ncc_impl(const pmxo_problem* pb, int view, int r, int c, double depth,
                int rotated, bool* clamp_flag, uint64_t* nsamples) {
    const pmxo_params& p = pb->p;
    const int hw = p.hw, Wn = 2 * hw + 1;
    const int Ws = rotated ? p.H : p.W, Hs = rotated ? p.W : p.H;
    const double scale = plane_scale(pb, r, c, rotated);
    const double* h1 = pb->H.data() + (size_t)view * 18;
    const double* h2 = h1 + 9;
    double T[9];
    for (int k = 0; k < 9; k++) T[k] = h1[k] - h2[k] / depth / scale;

    double sIp = 0.0, sIpIp = 0.0, sIIp = 0.0;
    int n = 0;
    for (int i = 0; i < Wn; i++) {
        const int rr = r - hw + i;
        if (rr < 0 || rr > Hs - 1) continue;
        double rs = 0.0, rss = 0.0, rx = 0.0;
        int j = 0;
        for (int k = 0; k < Wn; k++) {
            const int cc = c - hw + k;
            if (cc < 0 || cc > Ws - 1) continue;
            const double px = rotated ? rr + 0.5 : cc + 0.5;
            const double py = rotated ? cc + 0.5 : rr + 0.5;
            const double X = T[0] * px + T[1] * py + T[2];
            const double Y = T[3] * px + T[4] * py + T[5];
            const double Z = T[6] * px + T[7] * py + T[8];
            const double Ip = src_bilinear(pb, view, X / Z, Y / Z);
            const double Rv = ref_s(pb, rr, c - hw + j, rotated);
            j++;
            rss += Ip * Ip;
            rs += Ip;
            rx += Ip * Rv;
            n++;
        }
        sIp += rs;
        sIpIp += rss;
        sIIp += rx;
    }
    if (nsamples) *nsamples += (uint64_t)n;
    if (sIpIp == 0.0) return kSentinel;
    const int x = rotated ? r : c, y = rotated ? c : r;
    const double sI = pb->sum_I[(size_t)y * p.W + x];
    const double sII = pb->sum_II[(size_t)y * p.W + x];
    double c1 = sII - sI * sI / (double)n;
    double c2 = sIpIp - sIp * sIp / (double)n;
    if (clamp_flag && (near_floor(c1) || near_floor(c2))) *clamp_flag = true;
    c1 = c1 < kVarFloor ? 0.0 : c1;
    c2 = c2 < kVarFloor ? 0.0 : c2;
    const double den = std::sqrt(c1 * c2);
    if (den == 0.0) return 1.0;
    return 1.0 - (sIIp - sI * sIp / (double)n) / den;
}

struct ColumnRng {
    pmxo_rng* st;
    double uniform() { return pmxo_rng_uniform(st); }
};

/* computeMessageBackward (src/patchMatch.cu:669-690). */
void message_backward(double* m, const double* costs, int V, double alpha2,
                      ColumnRng& rng) {
    const double up = std::exp(-0.5 * kUniformX * kUniformX / alpha2);
    for (int v = 0; v < V; v++) {
        double e = costs[v];
        if (e == kSentinel) e = rng.uniform() * 2.0;
        e = std::exp(-0.5 * e * e / alpha2);
        const double z0 = m[v] * e * kFix + (1.0 - m[v]) * up * kChange;
        const double z1 = m[v] * e * kChange + (1.0 - m[v]) * up * kFix;
        m[v] = z0 / (z0 + z1);
    }
}

/* computeMessageForward (src/patchMatch.cu:692-715). */
void message_forward(double* m, const double* costs, int V, double alpha2,
                     ColumnRng& rng) {
    const double up = std::exp(-0.5 * kUniformX * kUniformX / alpha2);
    for (int v = 0; v < V; v++) {
        double e = costs[v];
        if (e == kSentinel) e = rng.uniform() * 2.0;
        e = std::exp(-0.5 * e * e / alpha2);
        const double z0 = (m[v] * kFix + (1.0 - m[v]) * kChange) * e;
        const double z1 = (m[v] * kChange + (1.0 - m[v]) * kFix) * up;
        m[v] = z0 / (z0 + z1);
    }
}

/* View draw (src/patchMatch.cu:940-961): the first view whose CDF entry is
 * >= u; the last view when none is (only possible with a NaN CDF). */
int pick_view(const double* cdf, int V, double u) {
    for (int i = 0; i < V; i++)
        if (u <= cdf[i]) return i;
    return V - 1;
}

} // namespace

extern "C" {

int pmxo_pick_view(const double* cdf, int V, double u) { return pick_view(cdf, V, u); }

/* Posterior of "view visible" along one column of L rows with the
 * reference's message recurrences (no sentinel costs): out[r] for r = 1..L-1
 * is the normalised product of the forward message through row r and the
 * backward message from rows r+1..L-1, as used by topToDown
 * (src/patchMatch.cu:858-916).  out[0] = NaN. */
void pmxo_hmm_posteriors(const double* costs, int L, double alpha, double* out) {
    const double alpha2 = alpha * alpha;
    pmxo_rng dummy{1, 0, 0, 0, 0, 0};
    ColumnRng rng{&dummy};
    std::vector<double> bwd(L);
    double m = 0.5;
    for (int r = L - 1; r >= 0; r--) {
        bwd[r] = m;
        message_backward(&m, &costs[r], 1, alpha2, rng);
    }
    m = 0.5;
    message_forward(&m, &costs[0], 1, alpha2, rng);
    out[0] = NAN;
    for (int r = 1; r < L; r++) {
        double f = m;
        message_forward(&f, &costs[r], 1, alpha2, rng);
        const double z0 = f * bwd[r], z1 = (1.0 - f) * (1.0 - bwd[r]);
        out[r] = z0 / (z0 + z1);
        m = f;
    }
}

double pmxo_ncc(const pmxo_problem* pb, int view, int row_s, int col_s,
                double depth, int rotated) {
    return ncc_impl(pb, view, row_s, col_s, depth, rotated, nullptr, nullptr);
}

/* computeAllCostGivenDepth (src/patchMatch.cu:775-816). */
void pmxo_cost_all_views(const pmxo_problem* pb, const double* depth_s,
                         double* cost_s, int rotated) {
    const pmxo_params& p = pb->p;
    const int Ws = rotated ? p.H : p.W, Hs = rotated ? p.W : p.H;
    uint64_t tot = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads()) reduction(+ : tot)
    for (int c = 0; c < Ws; c++) {
        uint64_t ns = 0;
        for (int r = 0; r < Hs; r++) {
            const double d = depth_s[(size_t)r * Ws + c];
            for (int v = 0; v < p.V; v++)
                cost_s[((size_t)v * Hs + r) * Ws + c] =
                    ncc_impl(pb, v, r, c, d, rotated, nullptr, &ns);
        }
        tot += ns;
    }
    pb->ncc_evals += (uint64_t)Ws * Hs * p.V;
    pb->samples += tot;
}

/* lowbias32 integer hash (C. Wellons, public domain constants). */
static uint32_t mix32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

/* The proposal hypothesis of pixel (x, y) in sweep k: inverse depth
 * 1/depth + spread h, h uniform in [-1, 1) from the hash of (y W + x, k),
 * clamped to [1/far, 1/near], all in FP32 with correctly rounded division
 * (the engine's arithmetic, reproduced bit for bit). */
static double proposal_draw(float depth, float spread, int x, int y, const pmxo_params& p,
                            int sweep_id) {
    const uint32_t k = mix32((uint32_t)(y * p.W + x) ^
                             mix32((uint32_t)sweep_id * 0x9E3779B9u + 0x632BE5ABu));
    const float h = (float)(k >> 8) * (1.0f / 8388608.0f) - 1.0f;
    const float lo = 1.0f / (float)p.far_d, hi = 1.0f / (float)p.near_d;
    const float u = std::fmin(std::fmax(std::fmaf(spread, h, 1.0f / depth), lo), hi);
    return (double)(1.0f / u);
}

/* topToDown (src/patchMatch.cu:818-1010) and downToTop (:1012-1209). */
void pmxo_sweep(const pmxo_problem* pb, double* depth_s, double* cost_s,
                pmxo_rng* states, int down, int rotated, int S,
                uint8_t* used_ids, const pmxo_trace* trace) {
    const pmxo_params& p = pb->p;
    const int V = p.V;
    const int Ws = rotated ? p.H : p.W, Hs = rotated ? p.W : p.H;
    const double alpha2 = p.alpha * p.alpha;
    uint64_t tot_evals = 0, tot_samples = 0;

    if (trace) {
        for (size_t i = 0; i < (size_t)Ws * Hs; i++) {
            if (trace->view_margin) trace->view_margin[i] = INFINITY;
            if (trace->cost_margin) trace->cost_margin[i] = INFINITY;
            if (trace->choice) trace->choice[i] = 255;
            if (trace->near_clamp) trace->near_clamp[i] = 0;
            if (trace->rand_depth) trace->rand_depth[i] = NAN;
            for (int v = 0; trace->cdf && v < V; v++) trace->cdf[i * V + v] = NAN;
            for (int q = 0; trace->cand_cost && q < 4; q++) trace->cand_cost[i * 4 + q] = NAN;
            if (trace->lpost_max) trace->lpost_max[i] = INFINITY;
        }
    }

#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads()) reduction(+ : tot_evals, tot_samples)
    for (int c = 0; c < Ws; c++) {
        /* Generator of the column: row 0 (plain) or column 0 (rotated) of the
         * original frame (src/patchMatch.cu:844-847). */
        ColumnRng rng{rotated ? &states[(size_t)c * p.W] : &states[(size_t)c]};
        std::vector<double> spm((size_t)V * Hs), m(V), costs(V), cdf(V);
        auto cost_at = [&](int v, int r) -> double& {
            return cost_s[((size_t)v * Hs + r) * Ws + c];
        };
        auto load_costs = [&](int r) {
            for (int v = 0; v < V; v++) costs[v] = cost_at(v, r);
        };
        uint64_t evals = 0, samples = 0;

        /* Backward messages, saved per row (src/patchMatch.cu:858-869 and
         * :1050-1061). */
        for (int v = 0; v < V; v++) m[v] = 0.5;
        const int first = down ? 0 : Hs - 1;
        for (int k = 0; k < Hs; k++) {
            const int r = down ? Hs - 1 - k : k;
            for (int v = 0; v < V; v++) spm[(size_t)v * Hs + r] = m[v];
            load_costs(r);
            message_backward(m.data(), costs.data(), V, alpha2, rng);
        }
        for (int v = 0; v < V; v++) spm[(size_t)v * Hs + first] = m[v];

        for (int v = 0; v < V; v++) m[v] = 0.5;
        load_costs(first);
        message_forward(m.data(), costs.data(), V, alpha2, rng);
        double d_former = depth_s[(size_t)first * Ws + c];

        for (int k = 1; k < Hs; k++) {
            const int r = down ? k : Hs - 1 - k;
            const int prev = down ? r - 1 : r + 1;
            std::vector<double> m_prev(m);
            for (int v = 0; v < V; v++) spm[(size_t)v * Hs + prev] = m[v];
            load_costs(r);
            message_forward(m.data(), costs.data(), V, alpha2, rng);
            /* posterior and its CDF (src/patchMatch.cu:909-923). */
            double lz0max = -INFINITY;
            for (int v = 0; v < V; v++) {
                const double b = spm[(size_t)v * Hs + r];
                const double z0 = m[v] * b;
                lz0max = std::fmax(lz0max, std::log(m[v]) + std::log(b));
                const double z1 = (1.0 - m[v]) * (1.0 - b);
                cdf[v] = z0 / (z0 + z1);
            }
            for (int v = 1; v < V; v++) cdf[v] += cdf[v - 1];
            const double total = cdf[V - 1];
            for (int v = 0; v < V; v++) cdf[v] /= total;

            const double d_cur = depth_s[(size_t)r * Ws + c];
            const double d_rand = rng.uniform() * (p.far_d - p.near_d) + p.near_d;
            /* a learned proposal (depth > 0) is a fourth candidate on the same
             * sampled views, compared last */
            const std::vector<double>& pv = rotated ? pb->prop_t : pb->prop;
            const std::vector<double>& sv = rotated ? pb->spread_t : pb->spread;
            const bool use_prop =
                !pv.empty() && (pb->prop_sweeps <= 0 || pb->sweep_id < pb->prop_sweeps);
            double d_prop = use_prop ? pv[(size_t)r * Ws + c] : 0.0;
            if (d_prop > 0.0 && sv[(size_t)r * Ws + c] > 0.0)
                d_prop = proposal_draw((float)d_prop, (float)sv[(size_t)r * Ws + c],
                                       rotated ? r : c, rotated ? c : r, p, pb->sweep_id);
            const int ncand = d_prop > 0.0 ? 4 : 3;
            double cost3[4] = {0.0, 0.0, 0.0, 0.0};
            double num_tested = 0.0;
            double vmargin = INFINITY;
            bool clamp_flag = false;
            for (int j = 0; j < S; j++) {
                const double u = rng.uniform();
                const int id = pick_view(cdf.data(), V, u);
                num_tested += 1.0;
                vmargin = std::fmin(vmargin, std::fabs(u - cdf[id]));
                if (id > 0) vmargin = std::fmin(vmargin, std::fabs(u - cdf[id - 1]));
                if (used_ids) used_ids[((size_t)j * Hs + r) * Ws + c] = (uint8_t)id;
                cost3[0] += ncc_impl(pb, id, r, c, d_former, rotated, &clamp_flag, &samples);
                cost3[1] += costs[id];
                cost3[2] += ncc_impl(pb, id, r, c, d_rand, rotated, &clamp_flag, &samples);
                evals += 2;
                if (ncand == 4) {
                    cost3[3] += ncc_impl(pb, id, r, c, d_prop, rotated, &clamp_flag, &samples);
                    evals += 1;
                }
            }
            const double inv = 1.0 / num_tested;
            for (int q = 0; q < 4; q++) cost3[q] *= inv;
            /* findMinCost (src/patchMatch.cu:613-618): strict <, ties to the
             * later candidate. */
            int idx = cost3[0] < cost3[1] ? 0 : 1;
            idx = cost3[idx] < cost3[2] ? idx : 2;
            if (ncand == 4) idx = cost3[idx] < cost3[3] ? idx : 3;
            const double cand[4] = {d_former, d_cur, d_rand, d_prop};
            const double best = cand[idx];
            depth_s[(size_t)r * Ws + c] = best;
            d_former = best;
            if (idx != 1) {
                for (int v = 0; v < V; v++)
                    cost_at(v, r) = ncc_impl(pb, v, r, c, best, rotated, &clamp_flag, &samples);
                evals += (uint64_t)V;
            }
            if (trace) {
                const size_t pi = (size_t)r * Ws + c;
                double cm = INFINITY;
                /* exact sentinel ties (every sampled view out of frame for
                 * both candidates) resolve the same way in any precision */
                for (int q = 0; q < ncand; q++)
                    if (q != idx && cand[q] != best &&
                        !(cost3[q] == kSentinel && cost3[idx] == kSentinel))
                        cm = std::fmin(cm, std::fabs(cost3[q] - cost3[idx]));
                if (trace->view_margin) trace->view_margin[pi] = vmargin;
                if (trace->cost_margin) trace->cost_margin[pi] = cm;
                if (trace->choice) trace->choice[pi] = (uint8_t)idx;
                if (trace->near_clamp) trace->near_clamp[pi] = clamp_flag ? 1 : 0;
                if (trace->rand_depth) trace->rand_depth[pi] = d_rand;
                if (trace->lpost_max) trace->lpost_max[pi] = lz0max;
                for (int v = 0; trace->cdf && v < V; v++) trace->cdf[pi * V + v] = cdf[v];
                for (int q = 0; trace->cand_cost && q < 4; q++)
                    trace->cand_cost[pi * 4 + q] = q < ncand ? cost3[q] : NAN;
            }
            m = m_prev;
            load_costs(r);
            message_forward(m.data(), costs.data(), V, alpha2, rng);
        }
        tot_evals += evals;
        tot_samples += samples;
    }
    pb->ncc_evals += tot_evals;
    pb->samples += tot_samples;
    pb->sweep_id++;
}

/* PatchMatch::run<W> (src/patchMatch.cu:307-407) including
 * transposeForward/transposeBackward (:247-261). */
void pmxo_run(const pmxo_problem* pb, double* depth, double* cost,
              pmxo_rng* states, int S, int iterations, uint8_t* used_ids) {
    const pmxo_params& p = pb->p;
    const int W = p.W, H = p.H, V = p.V;
    const size_t npx = (size_t)W * H;
    std::vector<double> depthT(npx), costT(npx * V);
    auto fwd = [&]() {
        pmxo_transpose(depth, W, H, depthT.data());
        for (int v = 0; v < V; v++)
            pmxo_transpose(cost + v * npx, W, H, costT.data() + v * npx);
    };
    auto bwd = [&]() {
        pmxo_transpose(depthT.data(), H, W, depth);
        for (int v = 0; v < V; v++)
            pmxo_transpose(costT.data() + v * npx, H, W, cost + v * npx);
    };
    for (int i = 0; i < iterations; i++) {
        const int s = (i == 0) ? 1 : S;
        fwd();
        if (i == 0) pmxo_cost_all_views(pb, depthT.data(), costT.data(), 1);
        pmxo_sweep(pb, depthT.data(), costT.data(), states, 1, 1, s, nullptr, nullptr);
        bwd();
        pmxo_sweep(pb, depth, cost, states, 1, 0, s, nullptr, nullptr);
        fwd();
        pmxo_sweep(pb, depthT.data(), costT.data(), states, 0, 1, s, nullptr, nullptr);
        bwd();
        pmxo_sweep(pb, depth, cost, states, 0, 0, s,
                   (i == iterations - 1) ? used_ids : nullptr, nullptr);
    }
}

void pmxo_problem_set_proposal(pmxo_problem* pb, const double* prop, const double* spread) {
    const int W = pb->p.W, H = pb->p.H;
    pb->sweep_id = 0;
    if (!prop) {
        pb->prop.clear();
        pb->prop_t.clear();
        pb->spread.clear();
        pb->spread_t.clear();
        return;
    }
    pb->prop.assign(prop, prop + (size_t)W * H);
    if (spread) pb->spread.assign(spread, spread + (size_t)W * H);
    else pb->spread.assign((size_t)W * H, 0.0);
    pb->prop_t.resize((size_t)W * H);
    pb->spread_t.resize((size_t)W * H);
    pmxo_transpose(pb->prop.data(), W, H, pb->prop_t.data());
    pmxo_transpose(pb->spread.data(), W, H, pb->spread_t.data());
}

void pmxo_problem_set_sweep_id(pmxo_problem* pb, int sweep_id) { pb->sweep_id = sweep_id; }

void pmxo_problem_set_proposal_sweeps(pmxo_problem* pb, int sweeps) { pb->prop_sweeps = sweeps; }

double pmxo_proposal_draw(float depth, float spread, int x, int y, int W, double near_d,
                          double far_d, int sweep_id) {
    pmxo_params p{};
    p.W = W;
    p.near_d = near_d;
    p.far_d = far_d;
    return spread > 0.0f ? proposal_draw(depth, spread, x, y, p, sweep_id) : (double)depth;
}

void pmxo_counters(const pmxo_problem* pb, uint64_t* ncc_evals, uint64_t* samples) {
    if (ncc_evals) *ncc_evals = pb->ncc_evals.load();
    if (samples) *samples = pb->samples.load();
}

void pmxo_set_threads(int n) { g_threads = n; }
int pmxo_get_threads(void) { return nthreads(); }

} // extern "C"
