/*
 * pmx_internal.h -- engine state and kernel launch interface shared by the
 * HIP kernels (pmx_kernels.hip) and the C ABI (pmx_capi.hip).
 *
 * Data layout in HBM (one engine = one reference view against V sources):
 *   sources   V layers, each (maxW + 4) x (maxH + 4) floats with a 2-pixel
 *             zero border; an image smaller than maxW x maxH is zero-padded
 *             like copyData (src/patchMatch.cu:72-95), so every bilinear tap
 *             is a plain load and "border = 0" needs no test.
 *   reference I (pre-blurred), sum I, sum I^2: W x H pitched float; per-pixel
 *             (mean, clamped n*variance) as float2, computed in FP64.
 *   state     depth (1 float) and V costs (V floats, view-innermost) per
 *             pixel, kept column-contiguous for the sweep frame in use:
 *             row-major for rotated (horizontal) sweeps, column-major for
 *             plain (vertical) sweeps, moved between the two by LDS-tiled
 *             transposes at the same points of the schedule as the
 *             reference's transposeForward/Backward (src/patchMatch.cu:247-261).
 *   messages  backward messages of the running sweep, V floats per pixel,
 *             column-contiguous.
 *   rng       one xorwow state per pixel (W x H, 24 bytes each); sweeps use
 *             row 0 (plain) or column 0 (rotated), as src/patchMatch.cu:844-847.
 */
#ifndef PMX_INTERNAL_H
#define PMX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/pmx.h"

#define PMX_SRC_PAD 2

struct pmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
};

/* Arguments of the sweep and cost kernels (device pointers). */
struct PmxSweepArgs {
    const float* src;      /* view 0, texel (0,0) of the padded source block */
    const void* src_block; /* start of the padded source block */
    long src_bytes;        /* its size (< 2^32: byte offsets are 32-bit) */
    long src_layer;        /* texels between layers */
    int src_pitch;         /* texels between rows */
    int src_h16;           /* 1: the block holds FP16 texels, 0: FP32 texel pairs */
    int srcW, srcH;        /* packed source size */
    const float* ref;      /* pre-blurred reference, W x H */
    int ref_pitch;         /* floats */
    const float2* stats;   /* (mean, clamped n*variance) per pixel, W x H */
    int stats_pitch;       /* float2 units */
    const double* Hm;      /* V x 18: H1 then H2, row-major */
    float* depth;          /* state: column c at depth + c*depth_ld */
    int depth_ld;
    float* cost;           /* state: column c at cost + c*cost_ld, [row][view] */
    int cost_ld;
    float* spm;            /* backward messages as log-odds: column c at spm + c*spm_ld, [row][view] */
    int spm_ld;
    pmx_rng_state* rng;    /* W x H */
    int rng_pitch;         /* bytes */
    uint8_t* ids;          /* S planes of W x H (original frame) or null */
    int ids_pitch;         /* bytes */
    unsigned long long* counters; /* [0] NCC evaluations, [1] window samples */
    const float* prop;     /* proposal depth, same layout as depth (0: none); or null */
    const float* prop_spread; /* its spread in inverse depth, same layout (0: exact) */
    int prop_ld;
    int sweep_id;          /* sweeps since prepare: keys the proposal draws */
    int W, H, V, hw, S;
    int Ws, Hs;            /* sweep frame */
    int down, rotated;
    float near_d, far_d;
    double alpha2;          /* SPMAlpha^2 in FP64 */
    double oX, oZ, iK0, iK1;
};

/* One batched sweep launch (pmx_sweep_batch_kernel): device tables of the
 * views' arguments and of the per-XCD column queues. */
struct PmxBatchLaunch {
    const PmxSweepArgs* tab; /* n views */
    const int* qstart;       /* 8 x (n + 1) */
    const int* qcol;         /* 8 x n */
    int n;
    unsigned* queue;         /* 8 counters, zero before the launch */
    int grid;                /* resident workgroups */
    int waves;               /* launch shape: 1 or 2 waves per column strip, or 3 */
    int hw, max_V, h16;
    int rot;                 /* the step's sweep frame (1: rotated) */
    int prp;                 /* every view carries a proposal */
    int sm;                  /* every view has at most 16 sources and 7 draws per row */
};

struct pmx_engine {
    pmx_ctx* ctx = nullptr;
    pmx_params p{};
    int W = 0, H = 0, V = 0, maxW = 0, maxH = 0;
    int image_format = PMX_IMG_F32;

    /* images */
    float* src_raw = nullptr;   /* padded source block as uploaded */
    float* src = nullptr;       /* padded source block, pre-blurred */
    uint16_t* src16 = nullptr;  /* the same block as FP16 (storage PMX_STORE_F16) */
    float2* src2 = nullptr;     /* the same block as FP32 texel pairs (t[x], t[x+1] - t[x])
                                 * (storage PMX_STORE_F32) */
    int storage = PMX_STORE_F32;
    long src_layer = 0;
    int src_pitch = 0;          /* floats */
    float* ref_raw = nullptr;   /* W x H, as uploaded */
    float* ref = nullptr;       /* W x H, pre-blurred (public _refImageData) */
    int ref_pitch = 0;          /* floats */
    float* sum_I = nullptr;
    float* sum_II = nullptr;
    float2* stats = nullptr;
    int stats_pitch = 0;        /* float2 units */

    double* Hm = nullptr;       /* V x 18 */
    std::vector<double> Hhost;
    pmx_camera ref_cam{};
    std::vector<pmx_camera> src_cams;
    std::vector<int> src_w, src_h;
    double invK[4] = {0, 0, 0, 0};

    /* state */
    float* depth_rm = nullptr;  int depth_rm_ld = 0;   /* row-major, ld in floats */
    float* depth_cm = nullptr;  int depth_cm_ld = 0;   /* column-major */
    float* cost_rm = nullptr;   int cost_rm_ld = 0;
    float* cost_cm = nullptr;   int cost_cm_ld = 0;
    float* spm = nullptr;       int spm_ld = 0;
    int layout_cm = 0;          /* 1: state currently in the column-major buffers */

    pmx_rng_state* rng = nullptr;
    int rng_pitch = 0;          /* bytes */
    uint8_t* used_ids = nullptr;
    int ids_pitch = 0;          /* bytes */
    int ids_layers = 0;

    /* reference-layout exports */
    float* out_depth = nullptr; int out_depth_pitch = 0;
    float* out_cost = nullptr;  int out_cost_pitch = 0;

    /* learned proposal: depth per pixel (0 = none), in both layouts */
    float* prop_rm = nullptr;   int prop_rm_ld = 0;
    float* prop_cm = nullptr;   int prop_cm_ld = 0;
    float* spread_rm = nullptr; /* proposal spread, same layouts and lds */
    float* spread_cm = nullptr;
    /* pmx_engine_propose's working buffers (network input, proposal planes),
     * kept from its first call to the engine's end: a device free waits for
     * the whole device, which would stall a batch running on another stream */
    void* prop_x = nullptr;
    float* prop_d = nullptr;
    bool has_prop = false;
    bool prop_seeds = true;     /* the proposal seeds the depth */
    int prop_sweeps = 0;        /* sweeps of a run that get the proposal (0: all) */
    int sweep_count = 0;        /* sweeps since the last prepare */

    unsigned long long* counters = nullptr;
    unsigned long long cost_all_evals = 0;   /* closed form, counted on the host */
    unsigned long long cost_all_samples = 0;
    bool prepared = false;
    int sweep_waves = 0;        /* wavefronts per sweep column (1, 2 or 4; 0 = automatic) */
    int num_cus = 0;            /* compute units of the device (queried once) */

    /* timing */
    bool timing = false;
    pmx_stats stats_acc{};
    std::vector<hipEvent_t> ev_pool;
};

/* ---- kernel launchers (pmx_kernels.hip); all return hipError_t ---- */
hipError_t pmx_launch_transpose_f32(hipStream_t s, const float* in, float* out, int width,
                                    int height, int ld_in, int ld_out);
hipError_t pmx_launch_transpose_rec(hipStream_t s, const float* in, float* out, int width,
                                    int height, int E, int ld_in, int ld_out);
hipError_t pmx_launch_rng_init(hipStream_t s, pmx_rng_state* st, int pitch_bytes, int width,
                               int height, int bx, int by, uint32_t seed);
hipError_t pmx_launch_fill_depth(hipStream_t s, pmx_rng_state* st, int st_pitch_bytes,
                                 float* out, int out_ld, int width, int height, float near_d,
                                 float far_d, int mode);
hipError_t pmx_launch_blur(hipStream_t s, float* data, int ld, int width, int height, int depth,
                           long layer_stride, const float* taps, int ntaps, float* tmp);
hipError_t pmx_launch_box_sums(hipStream_t s, const float* img, int img_ld, int width, int height,
                               int hw, float* out_I, float* out_sI, float* out_sII, int out_ld,
                               float2* stats, int stats_ld, double* tmp);
hipError_t pmx_launch_compose(hipStream_t s, const double* cams /* (V+1)*21 */, int V, double oX,
                              double oZ, double* H_out);
hipError_t pmx_launch_convert_u8(hipStream_t s, const uint8_t* in, int in_pitch, float* out,
                                 int out_ld, int width, int height);
hipError_t pmx_launch_cost_all(hipStream_t s, const PmxSweepArgs& a);
hipError_t pmx_launch_sweep(hipStream_t s, const PmxSweepArgs& a, int waves);
hipError_t pmx_launch_cost_volume(hipStream_t s, const PmxSweepArgs& a, int view0, int view1,
                                  uint16_t* out, int W0, int H0);
hipError_t pmx_launch_sweep_batch(hipStream_t s, const PmxBatchLaunch& b, int* per_cu);
hipError_t pmx_launch_zero_u32(hipStream_t s, unsigned* p, int n);
hipError_t pmx_launch_quantize_f16(hipStream_t s, float* data, uint16_t* out, size_t n);
hipError_t pmx_launch_pair_f32(hipStream_t s, const float* data, float2* out, size_t n);
hipError_t pmx_launch_seed_depth(hipStream_t s, const float* prop, int prop_ld, float* depth,
                                 int depth_ld, int width, int height);
hipError_t pmx_launch_gather_costs(hipStream_t s, const float* cost_rm, int cost_ld, float* out,
                                   int out_ld, int width, int height, int V);
hipError_t pmx_launch_scatter_costs(hipStream_t s, const float* in, int in_ld, float* cost_rm,
                                    int cost_ld, int width, int height, int V);

/* ---- learned proposal (pmx_cnn.hip); tensors NHWC bf16 stored as uint16 ---- */
hipError_t pmx_launch_conv_fwd(hipStream_t s, const uint16_t* X, const uint16_t* W,
                               const float* bias, uint16_t* Y, float* Yf, int N, int H, int Wd,
                               int Cin, int Cout, int stride, int relu);
hipError_t pmx_launch_conv_bwd_data(hipStream_t s, const uint16_t* dZ, const uint16_t* W2,
                                    uint16_t* dX, int N, int H, int Wd, int Cin, int Cout,
                                    int stride, int accumulate);
hipError_t pmx_launch_conv_bwd_weight(hipStream_t s, const uint16_t* X, const uint16_t* dZ,
                                      float* dW, int N, int H, int Wd, int Cin, int Cout,
                                      int stride);
hipError_t pmx_launch_bias_grad(hipStream_t s, const uint16_t* dZ, float* db, int P, int C);
hipError_t pmx_launch_relu_bwd(hipStream_t s, uint16_t* dY, const uint16_t* Y, size_t n);
hipError_t pmx_launch_weight_t(hipStream_t s, const uint16_t* W, uint16_t* W2, int Cout, int Cin);
hipError_t pmx_launch_upcat(hipStream_t s, const uint16_t* A, const uint16_t* S, uint16_t* out,
                            int N, int H, int W, int Ha, int Wa, int Ca, int Cs);
hipError_t pmx_launch_upcat_bwd(hipStream_t s, const uint16_t* dO, uint16_t* dA, uint16_t* dS,
                                int N, int H, int W, int Ha, int Wa, int Ca, int Cs);
hipError_t pmx_launch_head_loss(hipStream_t s, const float* Hf, const float* tgt, uint16_t* dH,
                                int P, int C, float lo, float hi, float inv_count, float w_conf,
                                float* sums);
hipError_t pmx_launch_adam(hipStream_t s, float* w, float* g, float* m, float* v, uint16_t* wb,
                           size_t n, size_t n_w, float lr, float b1, float b2, float eps, int t);
hipError_t pmx_launch_f32_to_bf16(hipStream_t s, const float* a, uint16_t* b, size_t n);
hipError_t pmx_launch_zero_f32(hipStream_t s, float* a, size_t n);
hipError_t pmx_launch_proposal(hipStream_t s, const float* Hf, int C, int H0, int W0, int W, int H,
                               float lo, float hi, float tau, float* out, int ld, float* spread,
                               int ld_spread, float* conf, int ld_conf);
hipError_t pmx_launch_target(hipStream_t s, const float* depth, int ld, int W, int H, int W0, int H0,
                             float* tgt);

/* The context a proposal network was created on (pmx_net.hip). */
pmx_ctx* pmx_net_context(const pmx_net* net);

/* Error reporting shared by the C ABI sources: records the message for
 * pmx_last_error and returns code. */
int pmx_fail(pmx_ctx* ctx, int code, const char* fmt, ...);

#endif
