/*
 * pmx_oracle.h -- scalar FP64 CPU restatement of the reference PatchMatch
 * multi-view-stereo kernels.  TEST INFRASTRUCTURE ONLY: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark load this
 * library as the checker.  The product path never links or calls it.
 *
 * Every function names the reference file:line it restates.  Arithmetic is
 * FP64 throughout; constants that the reference writes as FP32 literals
 * (0.999f, 1.0f-0.999f, 0.83f, 2.0f) are taken at their FP32 values and then
 * used in FP64, so the oracle evaluates the reference's own numbers.
 *
 * Parity status of each part (how it is pinned) is listed in the header of
 * pmx_oracle.cpp.
 */
#ifndef PMX_ORACLE_H
#define PMX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Project generator: Marsaglia xorwow (J. Stat. Softw. 8(14), 2003) with a
 * SplitMix64 seeding (Steele, Lea, Flood, OOPSLA 2014).  Replaces the
 * reference's curandState (src/array2d_psng.h:8-19). */
typedef struct pmxo_rng {
    uint32_t x, y, z, w, v, d;
} pmxo_rng;

void     pmxo_rng_seed(pmxo_rng* st, uint32_t id, uint32_t seed);
uint32_t pmxo_rng_next(pmxo_rng* st);
double   pmxo_rng_uniform(pmxo_rng* st); /* ((r>>8)+1) * 2^-24, in (0,1] */
/* Block-linear thread id of src/array2d_psng.cu:8-9 for a block_x*block_y grid. */
uint32_t pmxo_seed_id(int x, int y, int width, int block_x, int block_y);
/* setup_kernel (src/array2d_psng.cu:3-17): one state per pixel, row-major. */
void pmxo_rng_init_grid(pmxo_rng* states, int width, int height, int block_x,
                        int block_y, uint32_t seed);
/* generate_kernel_float (src/array2d_wrapper.cu:3-20).  mode 0: depth
 * (1-u)(far-near)+near (reference); mode 1: inverse depth uniform between
 * 1/far and 1/near. */
void pmxo_init_depth(pmxo_rng* states, int width, int height, double near_d,
                     double far_d, int mode, double* depth);

/* Pre-blur taps (src/GaussianBlurCUDA.cu:99-130).  mode 1: normalised
 * Gaussian; mode 2: the reference's uniform 1/width box.  Returns width. */
int  pmxo_blur_taps(double sigma, int mode, double* taps /* >= 45 */);
/* Separable filter, clamp-to-edge (src/GaussianBlurCUDA.cu:39-70). */
void pmxo_blur(const double* in, int width, int height, double sigma, int mode,
               double* out);
/* Array2d_refImg::filterImage (src/array2d_refImg.cu:11-131): window sums of
 * I and I^2 with zero outside the image. */
void pmxo_box_sums(const double* img, int width, int height, int hw,
                   double* sum_I, double* sum_II);
/* cudaTranspose::transpose2dData (src/cudaTranspose.h:8-71): out[x][y] = in[y][x]. */
void pmxo_transpose(const double* in, int width, int height, double* out);
/* Image::init_relative (src/image.cpp:4-13) in FP64.  Row-major 3x3 K, R;
 * C is the camera centre. */
void pmxo_compose(const double* Kr, const double* Rr, const double* Cr,
                  const double* Ks, const double* Rs, const double* Cs,
                  double oX, double oZ, double* H1, double* H2);

typedef struct pmxo_params {
    int    W, H;        /* reference size */
    int    V;           /* number of source views */
    int    srcW, srcH;  /* packed source block size (max over sources) */
    int    hw;          /* half window */
    double near_d, far_d;
    double alpha;       /* SPMAlpha */
    double oX, oZ;      /* plane orientation (oX, 0, oZ) */
    double invK[4];     /* reference K^-1 entries (0,0) (0,2) (1,1) (1,2) */
} pmxo_params;

/* Per-pixel record of how close each discrete decision came to flipping.
 * Arrays are in the sweep frame (Ws columns x Hs rows), row-major. */
typedef struct pmxo_trace {
    double*  view_margin;  /* min_j min_i |u_j - cdf_i| over the S draws */
    double*  cost_margin;  /* min |c_best - c_k| over candidates k with another depth */
    uint8_t* choice;       /* 0 former, 1 current, 2 random, 3 proposal; 255 = row not processed */
    uint8_t* near_clamp;   /* 1 if any variance term came within 0.1% of 1e-5 */
    /* Optional (NULL = not recorded); rows not processed hold NaN. */
    double*  cdf;          /* Ws*Hs*V: the normalised view CDF the row's draws read, [r][c][v] */
    double*  rand_depth;   /* Ws*Hs: the row's random depth hypothesis */
    double*  cand_cost;    /* Ws*Hs*4: averaged costs of former, current, random, proposal */
    /* Optional: log of the largest posterior numerator m_f * beta over the
     * row's views, as log m_f + log beta (+inf where not processed, -inf
     * where a message is exactly 0).  Between about -745 and -708 the FP64
     * numerators are denormal or round to 0 (every one 0: a NaN CDF), a
     * regime in which no other precision follows the view draws decision
     * for decision. */
    double*  lpost_max;
} pmxo_trace;

typedef struct pmxo_problem pmxo_problem;

/* ref: W*H row-major; srcs: V layers of srcW*srcH (zero outside each image);
 * H1H2: V*18 doubles (H1 row-major then H2). */
pmxo_problem* pmxo_problem_create(const pmxo_params* p, const double* ref,
                                  const double* srcs, const double* H1H2);
void   pmxo_problem_destroy(pmxo_problem* pb);
/* Box sums of the reference that the problem uses (W*H each). */
void   pmxo_problem_ref_sums(const pmxo_problem* pb, double* sum_I, double* sum_II);
/* computeNCC (src/patchMatch.cu:433-611) at sweep-frame pixel (row_s, col_s). */
double pmxo_ncc(const pmxo_problem* pb, int view, int row_s, int col_s,
                double depth, int rotated);
/* computeAllCostGivenDepth (src/patchMatch.cu:775-816) over a sweep frame.
 * depth_s: Ws*Hs; cost_s: V layers of Ws*Hs (layer v = rows v*Hs..). */
void   pmxo_cost_all_views(const pmxo_problem* pb, const double* depth_s,
                           double* cost_s, int rotated);
/* topToDown / downToTop (src/patchMatch.cu:818-1209) over a sweep frame.
 * states: W*H row-major (original frame).  used_ids: S layers of Ws*Hs or
 * NULL.  trace: NULL or arrays of Ws*Hs. */
void   pmxo_sweep(const pmxo_problem* pb, double* depth_s, double* cost_s,
                  pmxo_rng* states, int down, int rotated, int S,
                  uint8_t* used_ids, const pmxo_trace* trace);
/* PatchMatch::run<W> (src/patchMatch.cu:307-407).  depth: W*H; cost: V
 * layers of W*H; used_ids: S layers of W*H or NULL. */
void   pmxo_run(const pmxo_problem* pb, double* depth, double* cost,
                pmxo_rng* states, int S, int iterations, uint8_t* used_ids);
/* View draw of src/patchMatch.cu:940-961 (first i with u <= cdf[i]). */
int    pmxo_pick_view(const double* cdf, int V, double u);
/* Smoothing posterior of one view along a column of L costs, with the
 * message recurrences of src/patchMatch.cu:669-715 (no sentinel costs). */
void   pmxo_hmm_posteriors(const double* costs, int L, double alpha, double* out);
/* Learned proposal for the sweeps: W*H depths in the original frame (0 =
 * none at that pixel) and spreads in inverse depth (NULL = 0); NULL depths
 * remove it.  A sweep row adds the proposal as a fourth candidate after
 * former, current and random: the depth itself for spread 0, else a draw
 * keyed by the pixel and the sweep count (no reference counterpart: the
 * learned proposal of the MI355X engine).  Resets the sweep count to 0;
 * every pmxo_sweep advances it. */
void   pmxo_problem_set_proposal(pmxo_problem* pb, const double* prop, const double* spread);
void   pmxo_problem_set_sweep_id(pmxo_problem* pb, int sweep_id);
/* The proposal joins only the sweeps with count below `sweeps` (0: all), as
 * pmx_engine_set_proposal_sweeps of the engine. */
void   pmxo_problem_set_proposal_sweeps(pmxo_problem* pb, int sweeps);
/* The proposal hypothesis of pixel (x, y) of a W-wide image in sweep k. */
double pmxo_proposal_draw(float depth, float spread, int x, int y, int W, double near_d,
                          double far_d, int sweep_id);
/* Counters of the last calls (NCC evaluations and window samples). */
void   pmxo_counters(const pmxo_problem* pb, uint64_t* ncc_evals, uint64_t* samples);
void   pmxo_set_threads(int n);
int    pmxo_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
