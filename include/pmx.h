/*
 * pmx.h -- C ABI of the MI355X-native PatchMatch multi-view-stereo engine.
 *
 * Plain pointers and sizes only.  Pitches are in bytes (as the reference's
 * _pitchData, src/array2d_wrapper.h:17), device pointers come from
 * pmx_malloc_pitch / pmx_malloc, host pointers are borrowed for the call.
 * Every entry point returns PMX_OK (0) or an error code and never exits;
 * pmx_last_error() gives the message.  Work is ordered on the context's own
 * HIP stream; calls that return host data synchronise that stream.
 *
 * Each declaration cites the reference interface it replaces.
 */
#ifndef PMX_H
#define PMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMX_OK 0
#define PMX_ERR_ARG 1         /* invalid argument (reference: message + exit) */
#define PMX_ERR_HIP 2         /* HIP runtime error (reference: CUDA_SAFE_CALL exit) */
#define PMX_ERR_ALLOC 3       /* device allocation failed */
#define PMX_ERR_STATE 4       /* call not valid in the current engine state */
#define PMX_ERR_NO_DEVICE 5   /* no usable GPU */

#define PMX_BLUR_NONE 0
#define PMX_BLUR_GAUSS 1      /* normalised Gaussian taps */
#define PMX_BLUR_BOX 2        /* reference taps: uniform 1/width (src/GaussianBlurCUDA.cu:128-129) */

#define PMX_INIT_DEPTH 0      /* (1-u)(far-near)+near (src/array2d_wrapper.cu:15) */
#define PMX_INIT_INVERSE 1    /* inverse depth uniform in [1/far, 1/near] */

#define PMX_IMG_F32 0         /* float32 intensities */
#define PMX_IMG_U8 1          /* uchar, read as u/255 (cudaReadModeNormalizedFloat) */

#define PMX_STORE_F32 0       /* source and reference texels stored as float32 */
#define PMX_STORE_F16 1       /* stored as FP16; NCC sums still accumulate in FP32 */

#define PMX_MEMCPY_H2D 1
#define PMX_MEMCPY_D2H 2
#define PMX_MEMCPY_D2D 3

#define PMX_MAX_VIEWS 64
#define PMX_MAX_SAMPLES 64
#define PMX_MAX_HALF_WINDOW 26   /* src/patchMatch.cu:7 MAX_WINDOW_SIZE 53 */

typedef struct pmx_ctx pmx_ctx;
typedef struct pmx_engine pmx_engine;

/* Per-pixel generator state: Marsaglia xorwow, 24 bytes.  Replaces
 * curandState (src/array2d_psng.h:8). */
typedef struct pmx_rng_state {
    uint32_t x, y, z, w, v, d;
} pmx_rng_state;

/* Pinhole camera: row-major K and R, camera centre C (src/image.h:13-19). */
typedef struct pmx_camera {
    double K[9];
    double R[9];
    double C[3];
} pmx_camera;

/* Constructor arguments of PatchMatch (src/patchMatch.h:71-72) plus the
 * RNG layout fixed in src/main.cpp:115-116. */
typedef struct pmx_params {
    int half_window;        /* halfWindowSize, 1..22 (src/array2d_refImg.cu:72-97) */
    int num_samples;        /* numOfSamples (S) */
    int num_iterations;     /* numOfIterations */
    float depth_near;       /* nearRange */
    float depth_far;        /* farRange */
    float spm_alpha;        /* SPMAlpha */
    float gaussian_sigma;   /* gaussianSigma (pre-blur; unused by the reference) */
    int blur_mode;          /* PMX_BLUR_* */
    int init_mode;          /* PMX_INIT_* */
    float orientation_x;    /* orientationX */
    float orientation_z;    /* orientationZ */
    int block_x;            /* blockDim_x of the RNG seed layout (32) */
    int block_y;            /* blockDim_y of the RNG seed layout (16) */
    uint32_t rng_seed;      /* project seed mixed into every pixel's state */
} pmx_params;

/* Counters and stage timings of an engine (milliseconds, from HIP events on
 * the engine's stream when timing is enabled). */
typedef struct pmx_stats {
    uint64_t ncc_evals;       /* NCC evaluations of sweep kernels since the last reset */
    uint64_t ncc_samples;     /* their in-bounds window samples */
    uint64_t cost_all_evals;  /* NCC evaluations of the cost-all-views kernel */
    uint64_t cost_all_samples;
    int sweep_launches;
    int transpose_launches;
    double ms_sweep;          /* summed duration of sweep kernels */
    double ms_cost_all;       /* cost-all-views kernel */
    double ms_transpose;      /* state transposes */
    double ms_prepare;        /* blur + box sums + RNG init + depth init */
    double ms_sweep_max;      /* longest single sweep launch */
} pmx_stats;

/* ------------------------------------------------------------ context */
/* Replaces setBestGPUDevice (src/utility_CUDA.cpp:73-112) and the implicit
 * CUDA context: device < 0 picks device 0. */
int pmx_ctx_create(int device, pmx_ctx** out);
int pmx_ctx_destroy(pmx_ctx* ctx);
/* Message of the last failed call on ctx (or of the calling thread when ctx
 * is NULL).  Never NULL. */
const char* pmx_last_error(const pmx_ctx* ctx);
int pmx_ctx_sync(pmx_ctx* ctx);
/* The context's hipStream_t, for callers that order their own work after
 * the engine's. */
void* pmx_ctx_stream(pmx_ctx* ctx);
int pmx_device_count(int* count);

/* ------------------------------------------------------------- memory */
/* cudaMallocPitch in Array2D_wrapper (src/array2d_wrapper.h:23-24). */
int pmx_malloc_pitch(pmx_ctx* ctx, void** ptr, size_t* pitch_bytes,
                     size_t width_bytes, size_t height);
int pmx_malloc(pmx_ctx* ctx, void** ptr, size_t bytes);
int pmx_free(pmx_ctx* ctx, void* ptr);
/* Array2D_wrapper::copyData (src/array2d_wrapper.h:41-49). */
int pmx_memcpy2d(pmx_ctx* ctx, void* dst, size_t dpitch, const void* src,
                 size_t spitch, size_t width_bytes, size_t height, int kind);
int pmx_memset(pmx_ctx* ctx, void* ptr, int value, size_t bytes);

/* ------------------------------------------------- reference operators */
/* cudaTranspose::transpose2dData (src/cudaTranspose.h:52-71):
 * out[x][y] = in[y][x] for a width x height float plane. */
int pmx_transpose_f32(pmx_ctx* ctx, const float* in, float* out, int width,
                      int height, int pitch_in_bytes, int pitch_out_bytes);
/* Array2D_psng::randStateGen + setup_kernel (src/array2d_psng.cu:3-31):
 * one state per pixel, seeded by the block-linear id of a block_x*block_y
 * launch grid. */
int pmx_rng_init(pmx_ctx* ctx, pmx_rng_state* states, int pitch_bytes,
                 int width, int height, int block_x, int block_y, uint32_t seed);
/* Array2D_wrapper<float>::randNumGen + generate_kernel_float
 * (src/array2d_wrapper.h:137-145, src/array2d_wrapper.cu:3-20). */
int pmx_rng_fill_depth(pmx_ctx* ctx, pmx_rng_state* states, int state_pitch_bytes,
                       float* out, int out_pitch_bytes, int width, int height,
                       float depth_near, float depth_far, int init_mode);
/* Array2d_refImg::filterImage (src/array2d_refImg.cu:70-131): copies I and
 * writes the (2hw+1)^2 window sums of I and I^2 with zero border.  Sums are
 * accumulated in FP64 and stored as float. */
int pmx_box_sums(pmx_ctx* ctx, const float* img, int img_pitch_bytes, int width,
                 int height, int half_window, float* out_I, float* out_sum_I,
                 float* out_sum_II, int out_pitch_bytes);
/* GaussianBlurCUDA::FilterMultipleImages (src/GaussianBlurCUDA.cu:189-199):
 * in-place separable filter of `depth` stacked width x height planes
 * (plane k starts at row k*height), clamp-to-edge. */
int pmx_blur(pmx_ctx* ctx, float* data, int pitch_bytes, int width, int height,
             int depth, float sigma, int blur_mode);
/* pmx_blur without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The row pass goes to a
 * library-owned scratch of 2^23 floats per GPU, so width * height must be
 * at most 8388608 (PMX_ERR_ARG otherwise).  Stream-ordered on the context's
 * stream (no synchronisation; pmx_ctx_sync waits).  Every call on one GPU
 * shares that scratch: calls that overlap on one GPU (other contexts or
 * threads) race. */
int pmx_blur_noalloc(pmx_ctx* ctx, float* data, int pitch_bytes, int width, int height,
                     int depth, float sigma, int blur_mode);
/* Image::init_relative (src/image.cpp:4-13) for num_src views, composed in
 * FP64 on the GPU.  H_out: host, num_src * 18 doubles (H1 row-major, H2). */
int pmx_compose_homographies(pmx_ctx* ctx, const pmx_camera* ref,
                             const pmx_camera* srcs, int num_src, float orientation_x,
                             float orientation_z, double* H_out);

/* ------------------------------------------------------------- engine */
/* Argument checks of the PatchMatch constructor and runPatchMatch
 * (src/patchMatch.cu:137-142, :265-270; src/array2d_refImg.cu:96), without
 * touching a device.  pmx_engine_create runs the same checks. */
int pmx_validate_params(const pmx_params* params, int num_src);
/* PatchMatch::PatchMatch (src/patchMatch.cu:132-206).  ref_img and
 * src_imgs are host images (image_format PMX_IMG_F32 or PMX_IMG_U8) with
 * row pitches in bytes; sources may differ in size (packed like copyData,
 * src/patchMatch.cu:51-130).  The images stay resident on the device. */
int pmx_engine_create(pmx_ctx* ctx, const pmx_params* params, int image_format,
                      int ref_width, int ref_height, const void* ref_img,
                      int ref_pitch_bytes, int num_src, const int* src_widths,
                      const int* src_heights, const void* const* src_imgs,
                      const int* src_pitch_bytes, const pmx_camera* ref_cam,
                      const pmx_camera* src_cams, pmx_engine** out);
int pmx_engine_destroy(pmx_engine* eng);
/* Device bytes an engine of these sizes holds (pmx_engine_create, the
 * texel block of the storage (FP32 pairs or FP16) and the planes of a
 * proposal), without
 * touching a device: the admission measure of a caller that keeps many
 * views resident within an HBM budget.  No reference counterpart (the
 * reference sizes its buffers inline, src/patchMatch.cu:51-206). */
int pmx_engine_footprint(const pmx_params* params, int ref_width, int ref_height, int num_src,
                         const int* src_widths, const int* src_heights, int storage,
                         int with_proposal, size_t* bytes);
/* Constructor-time device work, re-runnable from the resident images:
 * pre-blur, reference box sums, RNG init, initial depth (src/patchMatch.cu:
 * 160-179). */
int pmx_engine_prepare(pmx_engine* eng);
/* PatchMatch::runPatchMatch / run<W> (src/patchMatch.cu:263-407). */
int pmx_engine_run(pmx_engine* eng);
/* Stages, for tests and tracing. */
int pmx_engine_cost_all_views(pmx_engine* eng, int rotated);            /* :338-343 */
int pmx_engine_sweep(pmx_engine* eng, int down, int rotated, int num_samples,
                     int record_ids);                                    /* :344-391 */
/* Host views of the engine state in the reference layouts (original frame):
 * depth W x H; costs V planes of W x H (src/patchMatch.cu:162); used ids S
 * planes of W x H (:170); rng W x H states. */
int pmx_engine_get_depth(pmx_engine* eng, float* host, int pitch_bytes);
int pmx_engine_set_depth(pmx_engine* eng, const float* host, int pitch_bytes);
int pmx_engine_get_costs(pmx_engine* eng, float* host);
int pmx_engine_set_costs(pmx_engine* eng, const float* host);
int pmx_engine_get_rng(pmx_engine* eng, pmx_rng_state* host);
int pmx_engine_set_rng(pmx_engine* eng, const pmx_rng_state* host);
int pmx_engine_get_used_ids(pmx_engine* eng, uint8_t* host);
/* Pre-blurred images the NCC reads: reference W x H, sources V planes of
 * maxW x maxH (zero outside each source).  Like the window sums below, they
 * exist from pmx_engine_prepare on (PMX_ERR_STATE before). */
int pmx_engine_get_images(pmx_engine* eng, float* ref_host, float* src_host);
int pmx_engine_get_ref_sums(pmx_engine* eng, float* sum_I, float* sum_II);
int pmx_engine_get_homographies(pmx_engine* eng, double* host);
int pmx_engine_get_dims(pmx_engine* eng, int* width, int* height, int* num_src,
                        int* max_width, int* max_height);
/* Device pointers of the output arrays in reference layout (valid after
 * pmx_engine_export): depth W x H pitched, costs W x (H*V), used ids
 * W x (H*S) uchar. */
int pmx_engine_export(pmx_engine* eng, float** depth, int* depth_pitch,
                      float** costs, int* cost_pitch, uint8_t** used_ids,
                      int* ids_pitch);
/* PatchMatch::_refImage (src/patchMatch.h:41, Array2d_refImg filled by the
 * constructor, src/patchMatch.cu:168-175) and _psngState (:49, seeded at
 * :164 and advanced by every sweep): device views of the engine's reference
 * image I (pre-blurred), its window sums of I and I^2 (W x H floats at
 * ref_pitch bytes) and its per-pixel generator states (W x H at rng_pitch
 * bytes).  Valid from pmx_engine_prepare until the engine is destroyed. */
int pmx_engine_export_state(pmx_engine* eng, float** ref, float** sum_I,
                            float** sum_II, int* ref_pitch, pmx_rng_state** rng,
                            int* rng_pitch);
int pmx_engine_set_timing(pmx_engine* eng, int enabled);
/* Launch shape of the sweep kernel: wavefronts per sweep column (1, 2 or
 * 4), or 0 = automatic, the default (two waves per column where the
 * columns fill at most half of the device's SIMDs, e.g. a 640x480 view;
 * one elsewhere, e.g. both sweep orientations of a 2048x1536 view).  No
 * reference counterpart (its sweeps use one 32-thread
 * block per 32 columns, src/patchMatch.cu:334).  Results do not depend on
 * it beyond floating-point summation order. */
int pmx_engine_set_sweep_waves(pmx_engine* eng, int waves);
int pmx_engine_get_sweep_waves(pmx_engine* eng, int* waves);
int pmx_engine_get_stats(pmx_engine* eng, pmx_stats* out);
int pmx_engine_reset_stats(pmx_engine* eng);

/* Texel storage of the pre-blurred images the NCC reads (PMX_STORE_*).  No
 * reference counterpart: the reference samples uchar textures normalised to
 * float (src/patchMatch.cu:27-28, :192-200).  PMX_STORE_F32 (the default)
 * keeps each blurred source texel with its right neighbour's difference
 * (8 bytes per texel, so a bilinear row blend is one FMA).  PMX_STORE_F16
 * rounds the blurred sources and reference to the nearest FP16 value (ties
 * to even), stores 2 bytes per texel and frees the FP32 pairs; homographies
 * stay FP64 and window sums FP32.  pmx_engine_get_images returns the
 * rounded values.  Takes effect at the next pmx_engine_prepare, which it
 * requires. */
int pmx_engine_set_storage(pmx_engine* eng, int storage);
int pmx_engine_get_storage(pmx_engine* eng, int* storage);

/* ------------------------------------------------------------ batches */
/* Many reference views (each its own engine: images, cameras, state) run
 * the runPatchMatch schedule (src/patchMatch.cu:263-407) together.  The
 * reference runs one refImageId per process (src/main.cpp:118-127); here
 * each of the 4 x numOfIterations sweep steps is ONE persistent launch over
 * the column strips of every view, longest columns first, and the whole
 * schedule (cost of all views, sweeps, transposes) is captured once in a
 * hipGraph and replayed.  The engines must share half_window, num_samples,
 * num_iterations and storage; their sizes and view counts may differ.
 * Each engine's results equal, bit for bit, a pmx_engine_run of that engine
 * alone with the same waves per column (launch shapes 1 and 3: one wave,
 * pmx_engine_set_sweep_waves(eng, 1); shape 2: two).  The batch borrows the
 * engines: destroy it before any of them. */
typedef struct pmx_batch pmx_batch;
int pmx_batch_create(pmx_ctx* ctx, pmx_engine* const* engines, int num_views, pmx_batch** out);
int pmx_batch_destroy(pmx_batch* batch);
/* pmx_engine_prepare of every engine. */
int pmx_batch_prepare(pmx_batch* batch);
/* The schedule of every view, on freshly prepared engines (PMX_ERR_STATE
 * otherwise).  ms (optional): wall time of the run from HIP events. */
int pmx_batch_run(pmx_batch* batch, float* ms);
/* Replay from a hipGraph (default 1); the graph is rebuilt when the launch
 * shape, storage, iterations or proposal presence change. */
int pmx_batch_set_graph(pmx_batch* batch, int enabled);
/* Launch shape: waves per column strip (1 or 2, with every register the
 * column code wants: two resident waves per SIMD), or 3 (one wave per strip
 * at three resident waves per SIMD, 168 registers with a few spilled; window
 * sizes 11 and 15, others run shape 1), and persistent workgroups (0 =
 * every slot the device keeps resident, else at least 8: one per XCD
 * queue). */
int pmx_batch_set_launch(pmx_batch* batch, int waves, int workgroups);
int pmx_batch_get_launch(pmx_batch* batch, int* waves, int* workgroups);
/* Autotuner: times full runs (prepare + run) for shapes 1, 2 and 3 at all,
 * 3/4 and 1/2 of the resident workgroups, keeps the fastest and returns it.
 * The engines are left run; prepare them again before pmx_batch_run. */
int pmx_batch_autotune(pmx_batch* batch, int* waves, int* workgroups, float* ms);

/* -------------------------------------------------- learned proposal */
/* The reference has no learned component.  A proposal is one extra depth
 * hypothesis per pixel per sweep, a fourth candidate beside the former,
 * current and random depths of src/patchMatch.cu:827-830 / :935-977; the
 * averaged NCC cost of findMinCost (:613-618) judges it like the others, and
 * pixels without a proposal run plain PatchMatch. */

/* Plane-sweep cost volume of the engine's (pre-blurred) images at 1/4
 * resolution, the network's input: x_dev is a device bf16 tensor of
 * ceil(H/4) x ceil(W/4) x 64 (NHWC); channel k < 32 is the mean 1 - NCC over
 * source views view0 and view1 of the window at full-resolution pixel
 * (min(4X+2, W-1), min(4Y+2, H-1)) on the fronto-parallel plane of inverse
 * depth 1/far + (1/near - 1/far)(k + 0.5)/32; channel 32 is that pixel's
 * reference intensity; channels 33..63 are 0.  Needs pmx_engine_prepare. */
int pmx_engine_cost_volume(pmx_engine* eng, int view0, int view1, void* x_dev);
/* Sets (depth_dev != NULL) or clears the per-pixel proposal: device W x H
 * float planes of depth (0 = no proposal at that pixel) and spread in
 * inverse depth (NULL = all 0).  In sweep k (counted from the last prepare)
 * the proposal hypothesis of pixel (x, y) is depth itself when the spread is
 * 0, else 1 / clamp(1/depth + spread * h, 1/far, 1/near) with h uniform in
 * [-1, 1) from an integer hash of (y W + x, k): the pixels' generators are
 * not touched.  Setting a proposal also seeds the current depth with it
 * where there is one; pmx_engine_prepare does so again after drawing the
 * initial depth. */
int pmx_engine_set_proposal(pmx_engine* eng, const float* depth_dev, const float* spread_dev,
                            int pitch_bytes);
/* The two source views nearest to the reference camera (shortest
 * baselines): the cost volume's default pair, the one the packaged network
 * was trained on. */
int pmx_engine_volume_views(pmx_engine* eng, int* view0, int* view1);
/* Whether a proposal also seeds the depth (default 1). */
int pmx_engine_set_proposal_seeding(pmx_engine* eng, int seed_depth);
/* The proposal hypothesis joins only the first `sweeps` sweeps of a run
 * (counted from the last prepare; 0, the default: every sweep).  No
 * reference counterpart: 4 limits it to the first iteration, where the
 * reference's depths are still its random initial ones. */
int pmx_engine_set_proposal_sweeps(pmx_engine* eng, int sweeps);
/* numOfIterations after construction (the proposal needs fewer). */
int pmx_engine_set_num_iterations(pmx_engine* eng, int iterations);

typedef struct pmx_net pmx_net;
/* The proposal network (4-level BF16 encoder-decoder, 3x3 convolutions with
 * 32/64/128/256 channels, MFMA kernels), sized for up to max_batch inputs of
 * H0 x W0 (the 1/4-resolution grid).  Its head predicts inverse depth in
 * [1/far, 1/near] and a confidence; weights are initialised from seed. */
int pmx_net_create(pmx_ctx* ctx, int max_batch, int H0, int W0, float depth_near,
                   float depth_far, unsigned long long seed, pmx_net** out);
int pmx_net_destroy(pmx_net* net);
/* Parameters are FP32 master values, all weights ([Cout][9][Cin] per layer,
 * 15 layers) then all biases; the convolutions read their bf16 rounding. */
int pmx_net_num_params(const pmx_net* net, size_t* n_total, size_t* n_weights);
int pmx_net_layer_offsets(const pmx_net* net, size_t* w_off, size_t* b_off); /* 15 each */
int pmx_net_get_params(pmx_net* net, float* host, size_t n);
int pmx_net_set_params(pmx_net* net, const float* host, size_t n);
int pmx_net_get_grads(pmx_net* net, float* host, size_t n);
int pmx_net_zero_grads(pmx_net* net);
/* Forward pass of N inputs (device bf16, N x H0 x W0 x 64). */
int pmx_net_forward(pmx_net* net, const void* x_dev, int N, int H0, int W0);
/* The last forward's head, N x H0 x W0 x 32 floats (channel 0: inverse-depth
 * logit, 1: confidence logit). */
int pmx_net_get_head(pmx_net* net, float* host, size_t n);
/* Gradients of the last forward's loss, added to the gradient buffer: L1 on
 * inverse depth u against tgt_dev (N x H0 x W0 floats, <= 0 = no target)
 * plus w_conf x L1 of the predicted error scale b against |u - t| (error
 * held fixed), both summed over pixels / (N H0 W0).  loss (host, 2 floats,
 * optional) receives the two mean terms. */
int pmx_net_backward(pmx_net* net, const void* x_dev, const float* tgt_dev, float w_conf,
                     float* loss);
/* Adam (beta1, beta2, eps) on the accumulated gradients, which it clears. */
int pmx_net_adam(pmx_net* net, float lr, float beta1, float beta2, float eps);
/* forward + backward + Adam(0.9, 0.999, 1e-8). */
int pmx_net_train_step(pmx_net* net, const void* x_dev, const float* tgt_dev, int N, int H0,
                       int W0, float lr, float w_conf, float* loss);
/* After a forward of batch 1: the head bilinearly upsampled to W x H.  The
 * confidence is P(|error| < 1% of u) = 1 - exp(-0.01 u / b); depth = 1/u
 * where it reaches tau, else 0 (no proposal); spread = 2 b.  spread_dev and
 * conf_dev are optional. */
int pmx_net_propose(pmx_net* net, int width, int height, float tau, float* depth_dev,
                    int depth_pitch_bytes, float* spread_dev, int spread_pitch_bytes,
                    float* conf_dev, int conf_pitch_bytes);
/* Cost volume -> forward -> proposal -> pmx_engine_set_proposal.  view0 or
 * view1 < 0 selects pmx_engine_volume_views.  Its working buffers (network
 * input, proposal planes) are allocated at the first call and kept until
 * pmx_engine_destroy (no device free while other streams run). */
int pmx_engine_propose(pmx_engine* eng, pmx_net* net, int view0, int view1, float tau);

/* Single operators of the network, on device bf16 NHWC tensors (channel
 * counts multiples of 32, pad 1, stride 1 or 2).  Wt: [Cout][9][Cin]. */
int pmx_cnn_conv_fwd(pmx_ctx* ctx, const void* X, const void* Wt, const float* bias, void* Y,
                     float* Yf, int N, int H, int W, int Cin, int Cout, int stride, int relu);
/* dX (+)= conv3x3^T(dZ); Wt2 = pmx_cnn_weight_t(Wt), [Cin][9][Cout]. */
int pmx_cnn_conv_bwd_data(pmx_ctx* ctx, const void* dZ, const void* Wt2, void* dX, int N, int H,
                          int W, int Cin, int Cout, int stride, int accumulate);
/* dW (FP32, [Cout][9][Cin]) += sum over output pixels of dZ x input patch. */
int pmx_cnn_conv_bwd_weight(pmx_ctx* ctx, const void* X, const void* dZ, float* dW, int N, int H,
                            int W, int Cin, int Cout, int stride);
int pmx_cnn_weight_t(pmx_ctx* ctx, const void* Wt, void* Wt2, int Cout, int Cin);
/* Training target: mean inverse depth of each 4x4 block of a W x H depth map. */
int pmx_cnn_target(pmx_ctx* ctx, const float* depth_dev, int depth_pitch_bytes, int width,
                   int height, float* tgt_dev);

#ifdef __cplusplus
}
#endif

#endif
