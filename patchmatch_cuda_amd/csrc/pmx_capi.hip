/*
 * pmx_capi.hip -- C ABI of the engine (declared in include/pmx.h).
 *
 * Owns the device buffers of one reference view and drives the schedule of
 * PatchMatch::run<W> (src/patchMatch.cu:307-407): per iteration a rotated
 * (left-to-right) sweep, a plain top-to-bottom sweep, a rotated
 * right-to-left sweep and a plain bottom-to-top sweep, with the state
 * transposed between the rotated and plain frames.  Unlike the reference it
 * allocates everything once (no malloc/free per sweep, src/patchMatch.cu:
 * 250-260, :337-391) and never synchronises between launches.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pmx_internal.h"

/* Scratch of pmx_blur_noalloc: the row pass of the separable filter writes
 * one width x height plane here, the column pass reads it back.  File-scope
 * device memory, one copy per GPU for the life of the process; host code
 * never passes this name to a kernel, only the address that
 * hipGetSymbolAddress returns for the current device (kept per device by
 * pmx_ctx_create). */
static constexpr size_t kBlurScratchFloats = (size_t)1 << 23; /* 32 MiB: a 4096 x 2048 plane */
static __device__ __align__(256) float scratch[kBlurScratchFloats];

namespace {

thread_local std::string g_thread_err;

int fail(pmx_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    g_thread_err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((ctx), PMX_ERR_HIP, "%s in %s at line %d", hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                               \
    } while (0)

#define PMXCHK(expr)             \
    do {                         \
        int rc_ = (expr);        \
        if (rc_ != PMX_OK) return rc_; \
    } while (0)

int dev_alloc(pmx_ctx* ctx, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(ctx, PMX_ERR_ALLOC, "device allocation of %zu bytes failed: %s", bytes,
                    hipGetErrorString(e));
    }
    return PMX_OK;
}

void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

/* Address of `scratch` on each device, set once per device by the first
 * context created on it (with that device current). */
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch[kMaxDevices] = {};

int scratch_setup(pmx_ctx* ctx) {
    if (ctx->device >= kMaxDevices)
        return fail(ctx, PMX_ERR_ARG, "device %d beyond the %d devices the library supports",
                    ctx->device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[ctx->device]) return PMX_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(ctx, PMX_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s",
                    ctx->device, hipGetErrorString(e));
    g_scratch[ctx->device] = (float*)p;
    return PMX_OK;
}

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

int blur_taps(float sigma, int mode, float* taps) {
    /* src/GaussianBlurCUDA.cu:99-130: width int(5 sigma), odd, in [5, 45]. */
    int width = (int)(5.0f * sigma);
    if (width % 2 == 0) width += 1;
    int sz = (width - 1) >> 1;
    if (width > 45) {
        sz = 22;
        width = 45;
    } else if (width < 5) {
        sz = 2;
        width = 5;
    }
    if (mode == PMX_BLUR_BOX) {
        for (int i = 0; i < width; i++) taps[i] = 1.0f / (float)width;
        return width;
    }
    double t[45], sum = 0.0;
    const double rv = -0.5 / ((double)sigma * (double)sigma);
    for (int i = -sz; i <= sz; i++) {
        t[i + sz] = std::exp((double)(i * i) * rv);
        sum += t[i + sz];
    }
    for (int i = 0; i < width; i++) taps[i] = (float)(t[i] / sum);
    return width;
}

/* ---- event timing ---- */
struct Span {
    int kind;
    hipEvent_t a, b;
};
enum { K_SWEEP = 0, K_COST_ALL, K_TRANSPOSE, K_PREPARE };

struct Timer {
    pmx_engine* e;
    std::vector<Span> spans;
    hipEvent_t get() {
        if (!e->ev_pool.empty()) {
            hipEvent_t ev = e->ev_pool.back();
            e->ev_pool.pop_back();
            return ev;
        }
        hipEvent_t ev = nullptr;
        if (hipEventCreate(&ev) != hipSuccess) return nullptr;
        return ev;
    }
    int begin(int kind) {
        if (!e->timing) return -1;
        Span s{kind, get(), get()};
        if (!s.a || !s.b) return -1;
        (void)hipEventRecord(s.a, e->ctx->stream);
        spans.push_back(s);
        return (int)spans.size() - 1;
    }
    void end(int id) {
        if (id < 0) return;
        (void)hipEventRecord(spans[id].b, e->ctx->stream);
    }
    void flush() {
        if (spans.empty()) return;
        (void)hipStreamSynchronize(e->ctx->stream);
        for (auto& s : spans) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, s.a, s.b);
            switch (s.kind) {
            case K_SWEEP:
                e->stats_acc.ms_sweep += ms;
                if (ms > e->stats_acc.ms_sweep_max) e->stats_acc.ms_sweep_max = ms;
                break;
            case K_COST_ALL: e->stats_acc.ms_cost_all += ms; break;
            case K_TRANSPOSE: e->stats_acc.ms_transpose += ms; break;
            default: e->stats_acc.ms_prepare += ms; break;
            }
            e->ev_pool.push_back(s.a);
            e->ev_pool.push_back(s.b);
        }
        spans.clear();
    }
};

PmxSweepArgs sweep_args(pmx_engine* e, int rotated) {
    PmxSweepArgs a{};
    a.src = e->src + (size_t)PMX_SRC_PAD * e->src_pitch + PMX_SRC_PAD;
    if (e->storage == PMX_STORE_F16) {
        a.src_block = e->src16;
        a.src_bytes = (long)sizeof(uint16_t) * e->src_layer * e->V;
        a.src_h16 = 1;
    } else {
        a.src_block = e->src2;
        a.src_bytes = (long)sizeof(float2) * e->src_layer * e->V;
        a.src_h16 = 0;
    }
    a.src_layer = e->src_layer;
    a.src_pitch = e->src_pitch;
    a.srcW = e->maxW;
    a.srcH = e->maxH;
    a.ref = e->ref;
    a.ref_pitch = e->ref_pitch;
    a.stats = e->stats;
    a.stats_pitch = e->stats_pitch;
    a.Hm = e->Hm;
    if (rotated) {
        a.depth = e->depth_rm;
        a.depth_ld = e->depth_rm_ld;
        a.cost = e->cost_rm;
        a.cost_ld = e->cost_rm_ld;
        a.Ws = e->H;
        a.Hs = e->W;
    } else {
        a.depth = e->depth_cm;
        a.depth_ld = e->depth_cm_ld;
        a.cost = e->cost_cm;
        a.cost_ld = e->cost_cm_ld;
        a.Ws = e->W;
        a.Hs = e->H;
    }
    a.spm = e->spm;
    a.spm_ld = a.Hs * e->V;
    a.rng = e->rng;
    a.rng_pitch = e->rng_pitch;
    a.ids = nullptr;
    a.ids_pitch = e->ids_pitch;
    a.counters = e->counters;
    a.prop = nullptr;
    a.prop_spread = nullptr;
    a.prop_ld = 0;
    a.sweep_id = e->sweep_count;
    if (e->has_prop && (e->prop_sweeps <= 0 || e->sweep_count < e->prop_sweeps)) {
        a.prop = rotated ? e->prop_rm : e->prop_cm;
        a.prop_spread = rotated ? e->spread_rm : e->spread_cm;
        a.prop_ld = rotated ? e->prop_rm_ld : e->prop_cm_ld;
    }
    a.W = e->W;
    a.H = e->H;
    a.V = e->V;
    a.hw = e->p.half_window;
    a.S = 1;
    a.rotated = rotated;
    a.down = 1;
    a.near_d = e->p.depth_near;
    a.far_d = e->p.depth_far;
    /* SPMAlphaSquare (src/patchMatch.cu:316), formed in FP64 */
    a.alpha2 = (double)e->p.spm_alpha * (double)e->p.spm_alpha;
    a.oX = e->p.orientation_x;
    a.oZ = e->p.orientation_z;
    a.iK0 = e->invK[0];
    a.iK1 = e->invK[1];
    return a;
}

/* Moves the state (depth + V costs) between the row-major and column-major
 * buffers. */
int to_layout(pmx_engine* e, int cm, Timer* tm, bool with_costs = true) {
    if (e->layout_cm == cm) return PMX_OK;
    pmx_ctx* ctx = e->ctx;
    const int id = tm ? tm->begin(K_TRANSPOSE) : -1;
    if (cm) {
        HIPCHK(ctx, pmx_launch_transpose_rec(ctx->stream, e->depth_rm, e->depth_cm, e->W, e->H, 1,
                                             e->depth_rm_ld, e->depth_cm_ld));
        if (with_costs)
            HIPCHK(ctx, pmx_launch_transpose_rec(ctx->stream, e->cost_rm, e->cost_cm, e->W, e->H,
                                                 e->V, e->cost_rm_ld, e->cost_cm_ld));
    } else {
        HIPCHK(ctx, pmx_launch_transpose_rec(ctx->stream, e->depth_cm, e->depth_rm, e->H, e->W, 1,
                                             e->depth_cm_ld, e->depth_rm_ld));
        if (with_costs)
            HIPCHK(ctx, pmx_launch_transpose_rec(ctx->stream, e->cost_cm, e->cost_rm, e->H, e->W,
                                                 e->V, e->cost_cm_ld, e->cost_rm_ld));
    }
    if (tm) tm->end(id);
    e->stats_acc.transpose_launches += with_costs ? 2 : 1;
    e->layout_cm = cm;
    return PMX_OK;
}

/* In-bounds window positions summed over the image: sum_x nx(x) * sum_y ny(y). */
static unsigned long long window_samples(int W, int H, int hw) {
    auto span = [hw](int n) {
        unsigned long long s = 0;
        for (int i = 0; i < n; i++) s += (unsigned long long)(std::min(i + hw, n - 1) - std::max(i - hw, 0) + 1);
        return s;
    };
    return span(W) * span(H);
}

int do_cost_all(pmx_engine* e, int rotated, Timer* tm) {
    PMXCHK(to_layout(e, 0, tm, false));
    PmxSweepArgs a = sweep_args(e, 1);
    a.rotated = rotated;
    const int id = tm ? tm->begin(K_COST_ALL) : -1;
    HIPCHK(e->ctx, pmx_launch_cost_all(e->ctx->stream, a));
    if (tm) tm->end(id);
    e->cost_all_evals += (unsigned long long)e->W * e->H * e->V;
    e->cost_all_samples += window_samples(e->W, e->H, e->p.half_window) * (unsigned long long)e->V;
    return PMX_OK;
}

/* Launch shape of one sweep: the engine's setting, or with the automatic
 * setting (0) two waves per column (at three waves per SIMD) when the
 * columns fill at most half of the SIMDs, else one.  A SIMD issues about
 * one instruction every 4 cycles over all of its waves, and a second wave
 * per column repeats the column's sequential logic: where the columns
 * already put one or two waves on every SIMD it only adds instructions
 * (1536-column sweeps of a 2048x1536 view: 16.9 ms per launch with one
 * wave per column, 17.7 ms with two).  Below that the second wave runs
 * on an otherwise idle SIMD and halves each row's hypothesis passes. */
int sweep_waves_for(pmx_engine* e, int Ws) {
    if (e->sweep_waves > 0) return e->sweep_waves;
    if (e->num_cus <= 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->ctx->device) !=
            hipSuccess)
            cus = 256;
        e->num_cus = cus;
    }
    return Ws <= 2 * e->num_cus ? 3 : 1;
}

int do_sweep(pmx_engine* e, int down, int rotated, int S, bool ids, Timer* tm) {
    PMXCHK(to_layout(e, rotated ? 0 : 1, tm));
    PmxSweepArgs a = sweep_args(e, rotated);
    a.down = down;
    a.S = S;
    if (ids) {
        if (S > e->ids_layers)
            return fail(e->ctx, PMX_ERR_ARG, "num_samples %d exceeds the %d used-id planes", S,
                        e->ids_layers);
        a.ids = e->used_ids;
    }
    const int id = tm ? tm->begin(K_SWEEP) : -1;
    HIPCHK(e->ctx, pmx_launch_sweep(e->ctx->stream, a, sweep_waves_for(e, a.Ws)));
    if (tm) tm->end(id);
    e->stats_acc.sweep_launches++;
    e->sweep_count++;
    return PMX_OK;
}

int check_engine(pmx_engine* e) {
    if (!e) return fail(nullptr, PMX_ERR_ARG, "engine is NULL");
    return PMX_OK;
}

int require_prepared(pmx_engine* e) {
    if (!e->prepared)
        return fail(e->ctx, PMX_ERR_STATE, "engine not prepared: call pmx_engine_prepare first");
    return PMX_OK;
}

} // namespace

int pmx_fail(pmx_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return fail(ctx, code, "%s", buf);
}

extern "C" {

/* ------------------------------------------------------------ context */

int pmx_device_count(int* count) {
    if (!count) return fail(nullptr, PMX_ERR_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return PMX_OK;
}

int pmx_ctx_create(int device, pmx_ctx** out) {
    if (!out) return fail(nullptr, PMX_ERR_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(nullptr, PMX_ERR_NO_DEVICE, "no HIP device is available");
    }
    if (device < 0) device = 0;
    if (device >= n)
        return fail(nullptr, PMX_ERR_ARG, "device %d out of range (%d devices)", device, n);
    pmx_ctx* ctx = new pmx_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, PMX_ERR_HIP, "context creation failed: %s", hipGetErrorString(e));
    }
    const int rc = scratch_setup(ctx);
    if (rc != PMX_OK) {
        const std::string msg = ctx->err;
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(nullptr, rc, "%s", msg.c_str());
    }
    *out = ctx;
    return PMX_OK;
}

int pmx_ctx_destroy(pmx_ctx* ctx) {
    if (!ctx) return PMX_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return PMX_OK;
}

const char* pmx_last_error(const pmx_ctx* ctx) {
    if (ctx) return ctx->err.c_str();
    return g_thread_err.c_str();
}

int pmx_ctx_sync(pmx_ctx* ctx) {
    if (!ctx) return fail(nullptr, PMX_ERR_ARG, "ctx is NULL");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return PMX_OK;
}

void* pmx_ctx_stream(pmx_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

/* ------------------------------------------------------------- memory */

int pmx_malloc_pitch(pmx_ctx* ctx, void** ptr, size_t* pitch, size_t width_bytes, size_t height) {
    if (!ctx || !ptr || !pitch) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_malloc_pitch");
    *ptr = nullptr;
    hipError_t e = hipMallocPitch(ptr, pitch, width_bytes ? width_bytes : 1, height ? height : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return fail(ctx, PMX_ERR_ALLOC, "hipMallocPitch(%zu x %zu) failed: %s", width_bytes, height,
                    hipGetErrorString(e));
    }
    return PMX_OK;
}

int pmx_malloc(pmx_ctx* ctx, void** ptr, size_t bytes) {
    if (!ctx || !ptr) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_malloc");
    return dev_alloc(ctx, ptr, bytes);
}

int pmx_free(pmx_ctx* ctx, void* ptr) {
    if (!ctx) return fail(nullptr, PMX_ERR_ARG, "ctx is NULL");
    if (ptr) HIPCHK(ctx, hipFree(ptr));
    return PMX_OK;
}

int pmx_memcpy2d(pmx_ctx* ctx, void* dst, size_t dpitch, const void* src, size_t spitch,
                 size_t width_bytes, size_t height, int kind) {
    if (!ctx || !dst || !src) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_memcpy2d");
    hipMemcpyKind k;
    switch (kind) {
    case PMX_MEMCPY_H2D: k = hipMemcpyHostToDevice; break;
    case PMX_MEMCPY_D2H: k = hipMemcpyDeviceToHost; break;
    case PMX_MEMCPY_D2D: k = hipMemcpyDeviceToDevice; break;
    default: return fail(ctx, PMX_ERR_ARG, "unknown copy kind %d", kind);
    }
    HIPCHK(ctx, hipMemcpy2DAsync(dst, dpitch, src, spitch, width_bytes, height, k, ctx->stream));
    if (kind != PMX_MEMCPY_D2D) HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return PMX_OK;
}

int pmx_memset(pmx_ctx* ctx, void* ptr, int value, size_t bytes) {
    if (!ctx || !ptr) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_memset");
    HIPCHK(ctx, hipMemsetAsync(ptr, value, bytes, ctx->stream));
    return PMX_OK;
}

/* ------------------------------------------------- reference operators */

int pmx_transpose_f32(pmx_ctx* ctx, const float* in, float* out, int width, int height,
                      int pitch_in, int pitch_out) {
    if (!ctx || !in || !out) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_transpose_f32");
    if (width <= 0 || height <= 0)
        return fail(ctx, PMX_ERR_ARG, "invalid transpose size %d x %d", width, height);
    if (pitch_in % 4 || pitch_out % 4 || pitch_in < width * 4 || pitch_out < height * 4)
        return fail(ctx, PMX_ERR_ARG, "invalid pitches %d / %d for %d x %d", pitch_in, pitch_out,
                    width, height);
    HIPCHK(ctx, pmx_launch_transpose_f32(ctx->stream, in, out, width, height, pitch_in / 4,
                                         pitch_out / 4));
    return PMX_OK;
}

int pmx_rng_init(pmx_ctx* ctx, pmx_rng_state* states, int pitch, int width, int height,
                 int block_x, int block_y, uint32_t seed) {
    if (!ctx || !states) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_rng_init");
    if (width <= 0 || height <= 0 || block_x <= 0 || block_y <= 0 ||
        pitch < width * (int)sizeof(pmx_rng_state))
        return fail(ctx, PMX_ERR_ARG, "invalid rng grid %d x %d (block %d x %d, pitch %d)", width,
                    height, block_x, block_y, pitch);
    HIPCHK(ctx, pmx_launch_rng_init(ctx->stream, states, pitch, width, height, block_x, block_y,
                                    seed));
    return PMX_OK;
}

int pmx_rng_fill_depth(pmx_ctx* ctx, pmx_rng_state* states, int state_pitch, float* out,
                       int out_pitch, int width, int height, float near_d, float far_d,
                       int init_mode) {
    if (!ctx || !states || !out) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_rng_fill_depth");
    if (width <= 0 || height <= 0 || out_pitch % 4 || out_pitch < width * 4)
        return fail(ctx, PMX_ERR_ARG, "invalid depth fill %d x %d pitch %d", width, height, out_pitch);
    if (!(near_d > 0.0f) || !(far_d > near_d))
        return fail(ctx, PMX_ERR_ARG, "invalid depth range [%g, %g]", near_d, far_d);
    HIPCHK(ctx, pmx_launch_fill_depth(ctx->stream, states, state_pitch, out, out_pitch / 4, width,
                                      height, near_d, far_d, init_mode));
    return PMX_OK;
}

int pmx_box_sums(pmx_ctx* ctx, const float* img, int img_pitch, int width, int height, int hw,
                 float* out_I, float* out_sI, float* out_sII, int out_pitch) {
    if (!ctx || !img) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_box_sums");
    if (hw < 1 || hw > 22)
        return fail(ctx, PMX_ERR_ARG, "Warning: too large size of windows (half window %d, 1..22)", hw);
    if (width <= 0 || height <= 0 || img_pitch % 4 || out_pitch % 4)
        return fail(ctx, PMX_ERR_ARG, "invalid box-sum geometry");
    void* tmp = nullptr;
    PMXCHK(dev_alloc(ctx, &tmp, sizeof(double) * 2 * (size_t)width * height));
    hipError_t e = pmx_launch_box_sums(ctx->stream, img, img_pitch / 4, width, height, hw, out_I,
                                       out_sI, out_sII, out_pitch / 4, nullptr, 0, (double*)tmp);
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    dev_free(tmp);
    HIPCHK(ctx, e);
    HIPCHK(ctx, e2);
    return PMX_OK;
}

int pmx_blur(pmx_ctx* ctx, float* data, int pitch, int width, int height, int depth, float sigma,
             int mode) {
    if (!ctx || !data) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_blur");
    if (mode == PMX_BLUR_NONE) return PMX_OK;
    if (mode != PMX_BLUR_GAUSS && mode != PMX_BLUR_BOX)
        return fail(ctx, PMX_ERR_ARG, "unknown blur mode %d", mode);
    if (!(sigma > 0.0f) || width <= 0 || height <= 0 || depth <= 0 || pitch % 4)
        return fail(ctx, PMX_ERR_ARG, "invalid blur arguments");
    float taps[45];
    const int n = blur_taps(sigma, mode, taps);
    void* tmp = nullptr;
    PMXCHK(dev_alloc(ctx, &tmp, sizeof(float) * (size_t)width * height));
    hipError_t e = pmx_launch_blur(ctx->stream, data, pitch / 4, width, height, depth,
                                   (long)(pitch / 4) * height, taps, n, (float*)tmp);
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    dev_free(tmp);
    HIPCHK(ctx, e);
    HIPCHK(ctx, e2);
    return PMX_OK;
}

int pmx_blur_noalloc(pmx_ctx* ctx, float* data, int pitch, int width, int height, int depth,
                     float sigma, int mode) {
    if (!ctx || !data) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_blur_noalloc");
    if (mode == PMX_BLUR_NONE) return PMX_OK;
    if (mode != PMX_BLUR_GAUSS && mode != PMX_BLUR_BOX)
        return fail(ctx, PMX_ERR_ARG, "unknown blur mode %d", mode);
    if (!(sigma > 0.0f) || width <= 0 || height <= 0 || depth <= 0 || pitch % 4 ||
        pitch < width * 4)
        return fail(ctx, PMX_ERR_ARG, "invalid blur arguments");
    if ((size_t)width * (size_t)height > kBlurScratchFloats)
        return fail(ctx, PMX_ERR_ARG,
                    "pmx_blur_noalloc: a %d x %d plane exceeds its scratch of %zu floats "
                    "(pmx_blur has no such limit)", width, height, kBlurScratchFloats);
    float* tmp = ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
    if (!tmp) return fail(ctx, PMX_ERR_STATE, "no scratch address for device %d", ctx->device);
    float taps[45];
    const int n = blur_taps(sigma, mode, taps);
    HIPCHK(ctx, pmx_launch_blur(ctx->stream, data, pitch / 4, width, height, depth,
                                (long)(pitch / 4) * height, taps, n, tmp));
    return PMX_OK;
}

int pmx_compose_homographies(pmx_ctx* ctx, const pmx_camera* ref, const pmx_camera* srcs,
                             int num_src, float oX, float oZ, double* H_out) {
    if (!ctx || !ref || !srcs || !H_out || num_src <= 0)
        return fail(ctx, PMX_ERR_ARG, "invalid arguments to pmx_compose_homographies");
    std::vector<double> cams((size_t)(num_src + 1) * 21);
    auto pack = [&](const pmx_camera& c, double* o) {
        std::memcpy(o, c.K, 9 * sizeof(double));
        std::memcpy(o + 9, c.R, 9 * sizeof(double));
        std::memcpy(o + 18, c.C, 3 * sizeof(double));
    };
    pack(*ref, cams.data());
    for (int v = 0; v < num_src; v++) pack(srcs[v], cams.data() + (size_t)(v + 1) * 21);
    void *dc = nullptr, *dh = nullptr;
    PMXCHK(dev_alloc(ctx, &dc, cams.size() * sizeof(double)));
    int rc = dev_alloc(ctx, &dh, (size_t)num_src * 18 * sizeof(double));
    if (rc != PMX_OK) {
        dev_free(dc);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(dc, cams.data(), cams.size() * sizeof(double),
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = pmx_launch_compose(ctx->stream, (const double*)dc, num_src, oX, oZ, (double*)dh);
    if (e == hipSuccess)
        e = hipMemcpyAsync(H_out, dh, (size_t)num_src * 18 * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(dc);
    dev_free(dh);
    HIPCHK(ctx, e);
    return PMX_OK;
}

/* ------------------------------------------------------------- engine */

static void engine_free(pmx_engine* e) {
    if (!e) return;
    void* bufs[] = {e->src_raw, e->src,     e->ref_raw,  e->ref,     e->sum_I,    e->sum_II,
                    e->stats,   e->Hm,      e->depth_rm, e->depth_cm, e->cost_rm, e->cost_cm,
                    e->spm,     e->rng,     e->used_ids, e->out_depth, e->out_cost, e->counters,
                    e->prop_rm, e->prop_cm, e->spread_rm, e->spread_cm, e->src16,
                    e->src2,    e->prop_x,  e->prop_d};
    for (void* b : bufs) dev_free(b);
    for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
    delete e;
}

static int validate(pmx_ctx* ctx, const pmx_params* params, int num_src) {
    if (!params) return fail(ctx, PMX_ERR_ARG, "params is NULL");
    /* src/patchMatch.cu:137-142 */
    if (num_src < 1) return fail(ctx, PMX_ERR_ARG, "Error: at least 2 images are needed for stereo");
    if (num_src > PMX_MAX_VIEWS)
        return fail(ctx, PMX_ERR_ARG, "at most %d source views are supported", PMX_MAX_VIEWS);
    const pmx_params& p = *params;
    /* src/patchMatch.cu:265-270 and src/array2d_refImg.cu:96 */
    if (p.half_window < 1 || p.half_window > 22)
        return fail(ctx, PMX_ERR_ARG,
                    "Warning: too large size of windows (half window %d, supported 1..22)",
                    p.half_window);
    if (p.num_samples < 1 || p.num_samples > PMX_MAX_SAMPLES)
        return fail(ctx, PMX_ERR_ARG, "numOfSamples %d outside 1..%d", p.num_samples, PMX_MAX_SAMPLES);
    if (p.num_iterations < 0) return fail(ctx, PMX_ERR_ARG, "numOfIterations must be >= 0");
    if (!(p.depth_near > 0.0f) || !(p.depth_far > p.depth_near))
        return fail(ctx, PMX_ERR_ARG, "invalid depth range [%g, %g]", p.depth_near, p.depth_far);
    if (!(p.spm_alpha > 0.0f)) return fail(ctx, PMX_ERR_ARG, "SPMAlpha must be positive");
    if (p.blur_mode != PMX_BLUR_NONE && p.blur_mode != PMX_BLUR_GAUSS && p.blur_mode != PMX_BLUR_BOX)
        return fail(ctx, PMX_ERR_ARG, "unknown blur mode %d", p.blur_mode);
    if (p.blur_mode != PMX_BLUR_NONE && !(p.gaussian_sigma > 0.0f))
        return fail(ctx, PMX_ERR_ARG, "gaussianSigma must be positive when blurring");
    if (p.init_mode != PMX_INIT_DEPTH && p.init_mode != PMX_INIT_INVERSE)
        return fail(ctx, PMX_ERR_ARG, "unknown init mode %d", p.init_mode);
    if (p.block_x < 1 || p.block_y < 1) return fail(ctx, PMX_ERR_ARG, "invalid RNG block layout");
    return PMX_OK;
}

int pmx_validate_params(const pmx_params* params, int num_src) {
    return validate(nullptr, params, num_src);
}

int pmx_engine_create(pmx_ctx* ctx, const pmx_params* params, int image_format, int ref_width,
                      int ref_height, const void* ref_img, int ref_pitch, int num_src,
                      const int* src_widths, const int* src_heights, const void* const* src_imgs,
                      const int* src_pitch, const pmx_camera* ref_cam, const pmx_camera* src_cams,
                      pmx_engine** out) {
    if (!ctx) return fail(nullptr, PMX_ERR_ARG, "ctx is NULL");
    if (!out || !params || !ref_img || !ref_cam)
        return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_engine_create");
    *out = nullptr;
    PMXCHK(validate(ctx, params, num_src));
    if (!src_widths || !src_heights || !src_imgs || !src_pitch || !src_cams)
        return fail(ctx, PMX_ERR_ARG, "NULL source arrays");
    const pmx_params& p = *params;
    if (image_format != PMX_IMG_F32 && image_format != PMX_IMG_U8)
        return fail(ctx, PMX_ERR_ARG, "unknown image format %d", image_format);
    const int esz = image_format == PMX_IMG_F32 ? 4 : 1;
    if (ref_width < 1 || ref_height < 1 || ref_pitch < ref_width * esz)
        return fail(ctx, PMX_ERR_ARG, "invalid reference image %d x %d pitch %d", ref_width,
                    ref_height, ref_pitch);
    int maxW = 0, maxH = 0;
    for (int v = 0; v < num_src; v++) {
        if (!src_imgs[v] || src_widths[v] < 1 || src_heights[v] < 1 ||
            src_pitch[v] < src_widths[v] * esz)
            return fail(ctx, PMX_ERR_ARG, "invalid source image %d", v);
        if (src_widths[v] > maxW) maxW = src_widths[v];
        if (src_heights[v] > maxH) maxH = src_heights[v];
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));

    pmx_engine* e = new pmx_engine();
    e->ctx = ctx;
    e->p = p;
    e->W = ref_width;
    e->H = ref_height;
    e->V = num_src;
    e->maxW = maxW;
    e->maxH = maxH;
    e->image_format = image_format;
    e->ref_cam = *ref_cam;
    e->src_cams.assign(src_cams, src_cams + num_src);
    e->src_w.assign(src_widths, src_widths + num_src);
    e->src_h.assign(src_heights, src_heights + num_src);
    const int W = e->W, H = e->H, V = e->V;
    hipStream_t s = ctx->stream;
    int rc = PMX_OK;
#define ALLOC(ptr, bytes)                                                   \
    do {                                                                    \
        void* q_ = nullptr;                                                 \
        rc = dev_alloc(ctx, &q_, (bytes));                                  \
        if (rc != PMX_OK) {                                                 \
            engine_free(e);                                                 \
            return rc;                                                      \
        }                                                                   \
        ptr = (decltype(ptr))q_;                                            \
    } while (0)
#define HIPCHK_E(expr)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            engine_free(e);                                                               \
            return fail(ctx, PMX_ERR_HIP, "%s in %s at line %d", hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                              \
        }                                                                                 \
    } while (0)

    /* sources: padded block */
    e->src_pitch = round_up(maxW + 2 * PMX_SRC_PAD, 64);
    e->src_layer = (long)e->src_pitch * (maxH + 2 * PMX_SRC_PAD);
    const size_t src_bytes = sizeof(float) * (size_t)e->src_layer * V;
    if (src_bytes >= (size_t)1 << 31) {
        engine_free(e);
        return fail(ctx, PMX_ERR_ARG,
                    "source block of %zu bytes exceeds the 2 GiB addressable by one engine",
                    src_bytes);
    }
    /* tap offsets are formed by a signed 24-bit multiply of the row index
     * and the row pitch in bytes (8 per texel pair) */
    if (e->src_pitch >= (1 << 20) || maxH >= (1 << 22)) {
        engine_free(e);
        return fail(ctx, PMX_ERR_ARG, "source images of %d x %d: at most 1048508 x 4194303 supported",
                    maxW, maxH);
    }
    ALLOC(e->src_raw, src_bytes);
    ALLOC(e->src, src_bytes);
    HIPCHK_E(hipMemsetAsync(e->src_raw, 0, src_bytes, s));
    e->ref_pitch = round_up(W, 64);
    const size_t plane = sizeof(float) * (size_t)e->ref_pitch * H;
    ALLOC(e->ref_raw, plane);
    ALLOC(e->ref, plane);
    ALLOC(e->sum_I, plane);
    ALLOC(e->sum_II, plane);
    e->stats_pitch = e->ref_pitch;
    ALLOC(e->stats, sizeof(float2) * (size_t)e->stats_pitch * H);

    /* upload images */
    {
        void* staging = nullptr;
        const size_t maxbytes = (size_t)std::max(W * H, maxW * maxH) * esz;
        if (image_format == PMX_IMG_U8) ALLOC(staging, maxbytes);
        auto upload = [&](const void* host, int hpitch, int w, int h, float* dst, int dld) -> hipError_t {
            if (image_format == PMX_IMG_F32)
                return hipMemcpy2DAsync(dst, (size_t)dld * 4, host, hpitch, (size_t)w * 4, h,
                                        hipMemcpyHostToDevice, s);
            hipError_t er = hipMemcpy2DAsync(staging, w, host, hpitch, w, h, hipMemcpyHostToDevice, s);
            if (er != hipSuccess) return er;
            er = pmx_launch_convert_u8(s, (const uint8_t*)staging, w, dst, dld, w, h);
            if (er != hipSuccess) return er;
            return hipStreamSynchronize(s);
        };
        hipError_t er = upload(ref_img, ref_pitch, W, H, e->ref_raw, e->ref_pitch);
        for (int v = 0; v < V && er == hipSuccess; v++)
            er = upload(src_imgs[v], src_pitch[v], src_widths[v], src_heights[v],
                        e->src_raw + (size_t)v * e->src_layer + (size_t)PMX_SRC_PAD * e->src_pitch +
                            PMX_SRC_PAD,
                        e->src_pitch);
        if (er == hipSuccess) er = hipStreamSynchronize(s);
        dev_free(staging);
        HIPCHK_E(er);
    }

    /* state */
    e->depth_rm_ld = round_up(W, 64);
    e->depth_cm_ld = round_up(H, 64);
    e->cost_rm_ld = round_up(W * V, 64);
    e->cost_cm_ld = round_up(H * V, 64);
    ALLOC(e->depth_rm, sizeof(float) * (size_t)e->depth_rm_ld * H);
    ALLOC(e->depth_cm, sizeof(float) * (size_t)e->depth_cm_ld * W);
    ALLOC(e->cost_rm, sizeof(float) * (size_t)e->cost_rm_ld * H);
    ALLOC(e->cost_cm, sizeof(float) * (size_t)e->cost_cm_ld * W);
    /* messages; also the scratch plane of the pre-blur */
    ALLOC(e->spm, std::max(sizeof(float) * (size_t)W * H * V, sizeof(float) * (size_t)maxW * maxH));
    HIPCHK_E(hipMemsetAsync(e->cost_rm, 0, sizeof(float) * (size_t)e->cost_rm_ld * H, s));
    e->rng_pitch = (int)(sizeof(pmx_rng_state) * (size_t)W);
    ALLOC(e->rng, (size_t)e->rng_pitch * H);
    e->ids_layers = p.num_samples;
    e->ids_pitch = round_up(W, 64);
    ALLOC(e->used_ids, (size_t)e->ids_pitch * H * e->ids_layers);
    HIPCHK_E(hipMemsetAsync(e->used_ids, 0, (size_t)e->ids_pitch * H * e->ids_layers, s));
    e->out_depth_pitch = (int)sizeof(float) * e->depth_rm_ld;
    e->out_cost_pitch = (int)sizeof(float) * e->ref_pitch;
    ALLOC(e->out_cost, (size_t)e->out_cost_pitch * H * V);
    ALLOC(e->counters, sizeof(unsigned long long) * 4);
    HIPCHK_E(hipMemsetAsync(e->counters, 0, sizeof(unsigned long long) * 4, s));

    /* homographies: FP64 on the device */
    {
        ALLOC(e->Hm, sizeof(double) * 18 * V);
        std::vector<double> cams((size_t)(V + 1) * 21);
        auto pack = [&](const pmx_camera& c, double* o) {
            std::memcpy(o, c.K, 9 * sizeof(double));
            std::memcpy(o + 9, c.R, 9 * sizeof(double));
            std::memcpy(o + 18, c.C, 3 * sizeof(double));
        };
        pack(*ref_cam, cams.data());
        for (int v = 0; v < V; v++) pack(src_cams[v], cams.data() + (size_t)(v + 1) * 21);
        /* the cameras pass through the message plane, not yet in use (no
         * temporary allocation: a device free waits for the whole device) */
        double* dc = (double*)e->spm;
        hipError_t er = hipMemcpyAsync(dc, cams.data(), cams.size() * sizeof(double),
                                       hipMemcpyHostToDevice, s);
        if (er == hipSuccess)
            er = pmx_launch_compose(s, dc, V, p.orientation_x, p.orientation_z, e->Hm);
        e->Hhost.resize((size_t)V * 18);
        if (er == hipSuccess)
            er = hipMemcpyAsync(e->Hhost.data(), e->Hm, sizeof(double) * 18 * V,
                                hipMemcpyDeviceToHost, s);
        if (er == hipSuccess) er = hipStreamSynchronize(s);
        HIPCHK_E(er);
        /* K^-1 entries (0,0) (0,2) (1,1) (1,2) of the reference
         * (src/patchMatch.cu:154-157), FP64 */
        const double* K = ref_cam->K;
        const double a = K[0], b = K[1], c = K[2], d = K[3], ee = K[4], f = K[5], g = K[6],
                     h = K[7], i = K[8];
        const double A = ee * i - f * h, B = -(d * i - f * g), C = d * h - ee * g;
        const double det = a * A + b * B + c * C;
        if (!(std::fabs(det) > 0.0)) {
            engine_free(e);
            return fail(ctx, PMX_ERR_ARG, "reference K is singular");
        }
        const double r = 1.0 / det;
        e->invK[0] = A * r;
        e->invK[1] = (b * f - c * ee) * r;
        e->invK[2] = (a * i - c * g) * r;
        e->invK[3] = -(a * f - c * d) * r;
    }

    HIPCHK_E(hipStreamSynchronize(s));
#undef ALLOC
#undef HIPCHK_E
    *out = e;
    return PMX_OK;
}

int pmx_engine_footprint(const pmx_params* params, int ref_width, int ref_height, int num_src,
                         const int* src_widths, const int* src_heights, int storage,
                         int with_proposal, size_t* bytes) {
    if (!params || !bytes || !src_widths || !src_heights || num_src < 1 || ref_width < 1 ||
        ref_height < 1)
        return fail(nullptr, PMX_ERR_ARG, "invalid argument to pmx_engine_footprint");
    int maxW = 0, maxH = 0;
    for (int v = 0; v < num_src; v++) {
        maxW = std::max(maxW, src_widths[v]);
        maxH = std::max(maxH, src_heights[v]);
    }
    /* the allocations of pmx_engine_create, pmx_engine_set_storage and
     * pmx_engine_set_proposal, each rounded up to hipMalloc's 256-byte granule */
    auto g = [](size_t b) { return (std::max<size_t>(b, 256) + 255) / 256 * 256; };
    const size_t W = (size_t)ref_width, H = (size_t)ref_height, V = (size_t)num_src;
    const size_t src_pitch = (size_t)round_up(maxW + 2 * PMX_SRC_PAD, 64);
    const size_t src_layer = src_pitch * (size_t)(maxH + 2 * PMX_SRC_PAD);
    const size_t ref_pitch = (size_t)round_up(ref_width, 64);
    size_t t = 0;
    t += 2 * g(sizeof(float) * src_layer * V);                       /* src_raw, src */
    t += 4 * g(sizeof(float) * ref_pitch * H);                        /* ref_raw, ref, sums */
    t += g(sizeof(float2) * ref_pitch * H);                           /* stats */
    t += g(sizeof(double) * 18 * V) + g(sizeof(double) * (V + 1) * 21);
    t += g(sizeof(float) * (size_t)round_up(ref_width, 64) * H);      /* depth_rm */
    t += g(sizeof(float) * (size_t)round_up(ref_height, 64) * W);     /* depth_cm */
    t += g(sizeof(float) * (size_t)round_up(ref_width * num_src, 64) * H);
    t += g(sizeof(float) * (size_t)round_up(ref_height * num_src, 64) * W);
    t += g(std::max(sizeof(float) * W * H * V, sizeof(float) * (size_t)maxW * maxH));
    t += g(sizeof(pmx_rng_state) * W * H);
    t += g((size_t)round_up(ref_width, 64) * H * (size_t)params->num_samples);
    t += g(sizeof(float) * ref_pitch * H * V);                        /* out_cost */
    t += g(sizeof(unsigned long long) * 4);
    if (storage == PMX_STORE_F16)
        t += g(sizeof(uint16_t) * src_layer * V);
    else
        t += g(sizeof(float2) * src_layer * V);
    if (with_proposal) {
        t += 2 * g(sizeof(float) * (size_t)round_up(ref_width, 64) * H);
        t += 2 * g(sizeof(float) * (size_t)round_up(ref_height, 64) * W);
        /* pmx_engine_propose's network input and proposal planes */
        t += g(sizeof(uint16_t) * (size_t)((ref_width + 3) / 4) * ((ref_height + 3) / 4) * 64);
        t += g(2 * sizeof(float) * W * H);
    }
    *bytes = t;
    return PMX_OK;
}

int pmx_engine_destroy(pmx_engine* e) {
    if (!e) return PMX_OK;
    if (e->ctx && e->ctx->stream) (void)hipStreamSynchronize(e->ctx->stream);
    engine_free(e);
    return PMX_OK;
}

static int prepare_impl(pmx_engine* e, Timer* tm) {
    pmx_ctx* ctx = e->ctx;
    hipStream_t s = ctx->stream;
    const int W = e->W, H = e->H, V = e->V;
    const int id = tm ? tm->begin(K_PREPARE) : -1;
    /* pre-blur: the reference and every source over its own extent; the
     * padded block stays zero outside each image */
    const size_t src_bytes = sizeof(float) * (size_t)e->src_layer * V;
    HIPCHK(ctx, hipMemcpyAsync(e->src, e->src_raw, src_bytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(e->ref, e->ref_raw, sizeof(float) * (size_t)e->ref_pitch * H,
                               hipMemcpyDeviceToDevice, s));
    if (e->p.blur_mode != PMX_BLUR_NONE) {
        float taps[45];
        const int n = blur_taps(e->p.gaussian_sigma, e->p.blur_mode, taps);
        float* tmp = (float*)e->spm; /* >= max(W*H, maxW*maxH) floats */
        HIPCHK(ctx, pmx_launch_blur(s, e->ref, e->ref_pitch, W, H, 1, 0, taps, n, tmp));
        for (int v = 0; v < V; v++) {
            const int w = e->src_w[v], h = e->src_h[v];
            float* layer = e->src + (size_t)v * e->src_layer + (size_t)PMX_SRC_PAD * e->src_pitch + PMX_SRC_PAD;
            HIPCHK(ctx, pmx_launch_blur(s, layer, e->src_pitch, w, h, 1, 0, taps, n, tmp));
        }
    }
    /* FP16 storage: the sources the NCC reads become FP16 (the FP32 block
     * keeps the same rounded values for readback), and the reference is
     * rounded the same way before its window sums */
    if (e->storage == PMX_STORE_F16) {
        HIPCHK(ctx, pmx_launch_quantize_f16(s, e->src, e->src16, (size_t)e->src_layer * V));
        HIPCHK(ctx, pmx_launch_quantize_f16(s, e->ref, nullptr, (size_t)e->ref_pitch * H));
    } else {
        /* FP32 storage: each texel with its right neighbour's difference,
         * so a bilinear row blend is one FMA (the pair block, under 4 GiB,
         * is allocated at the first FP32 prepare: engines switched to FP16
         * before preparing never hold it) */
        if (!e->src2) PMXCHK(dev_alloc(ctx, (void**)&e->src2, 2 * src_bytes));
        HIPCHK(ctx, pmx_launch_pair_f32(s, e->src, e->src2, (size_t)e->src_layer * V));
    }
    /* reference box sums and NCC statistics (FP64) */
    double* dtmp = (double*)e->cost_cm; /* >= 2*W*H doubles when V >= 4; else allocate */
    void* own = nullptr;
    if ((size_t)e->cost_cm_ld * W * sizeof(float) < sizeof(double) * 2 * (size_t)W * H) {
        PMXCHK(dev_alloc(ctx, &own, sizeof(double) * 2 * (size_t)W * H));
        dtmp = (double*)own;
    }
    hipError_t er = pmx_launch_box_sums(s, e->ref, e->ref_pitch, W, H, e->p.half_window, nullptr,
                                        e->sum_I, e->sum_II, e->ref_pitch, e->stats, e->stats_pitch,
                                        dtmp);
    if (own) {
        hipError_t e2 = hipStreamSynchronize(s);
        dev_free(own);
        if (er == hipSuccess) er = e2;
    }
    HIPCHK(ctx, er);
    /* RNG + initial depth (src/patchMatch.cu:164-166) */
    HIPCHK(ctx, pmx_launch_rng_init(s, e->rng, e->rng_pitch, W, H, e->p.block_x, e->p.block_y,
                                    e->p.rng_seed));
    HIPCHK(ctx, pmx_launch_fill_depth(s, e->rng, e->rng_pitch, e->depth_rm, e->depth_rm_ld, W, H,
                                      e->p.depth_near, e->p.depth_far, e->p.init_mode));
    /* a learned proposal seeds the depth where it has one (the draws above
     * keep every pixel's generator in step with the plain method) */
    if (e->has_prop && e->prop_seeds)
        HIPCHK(ctx, pmx_launch_seed_depth(s, e->prop_rm, e->prop_rm_ld, e->depth_rm, e->depth_rm_ld,
                                          W, H));
    if (tm) tm->end(id);
    e->layout_cm = 0;
    e->prepared = true;
    e->sweep_count = 0;
    return PMX_OK;
}

int pmx_engine_prepare(pmx_engine* e) {
    PMXCHK(check_engine(e));
    Timer tm{e, {}};
    int rc = prepare_impl(e, &tm);
    tm.flush();
    return rc;
}

int pmx_engine_run(pmx_engine* e) {
    PMXCHK(check_engine(e));
    PMXCHK(require_prepared(e));
    Timer tm{e, {}};
    const int iters = e->p.num_iterations;
    int rc = PMX_OK;
    for (int i = 0; i < iters && rc == PMX_OK; i++) {
        const int S = (i == 0) ? 1 : e->p.num_samples; /* src/patchMatch.cu:328-331 */
        if (i == 0) rc = do_cost_all(e, 1, &tm);
        if (rc == PMX_OK) rc = do_sweep(e, 1, 1, S, false, &tm);   /* left to right */
        if (rc == PMX_OK) rc = do_sweep(e, 1, 0, S, false, &tm);   /* top to bottom */
        if (rc == PMX_OK) rc = do_sweep(e, 0, 1, S, false, &tm);   /* right to left */
        if (rc == PMX_OK) rc = do_sweep(e, 0, 0, S, i == iters - 1, &tm); /* bottom to top */
    }
    if (rc == PMX_OK) rc = to_layout(e, 0, &tm);
    tm.flush();
    if (rc == PMX_OK) HIPCHK(e->ctx, hipGetLastError());
    return rc;
}

/* The two source views whose centres are nearest to the reference's (the
 * shortest baselines), for the cost volume: with 32 planes over the whole
 * inverse-depth range, a plane step moves a pixel by f b du, so only short
 * baselines keep neighbouring planes within the NCC's capture range (on a
 * 2048x1536 view with 10 sources the widest pair put 13% of the volume's
 * minima within 5% of the true inverse depth, the nearest 36%). */
static void volume_views(const pmx_engine* e, int* v0, int* v1) {
    std::vector<std::pair<double, int>> b;
    for (int v = 0; v < e->V; v++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) {
            const double d = e->src_cams[v].C[k] - e->ref_cam.C[k];
            s += d * d;
        }
        b.push_back({s, v});
    }
    std::sort(b.begin(), b.end());
    *v0 = b[0].second;
    *v1 = b.size() > 1 ? b[1].second : b[0].second;
}

int pmx_engine_volume_views(pmx_engine* e, int* view0, int* view1) {
    PMXCHK(check_engine(e));
    if (!view0 || !view1) return fail(e->ctx, PMX_ERR_ARG, "NULL argument");
    volume_views(e, view0, view1);
    return PMX_OK;
}

int pmx_engine_set_storage(pmx_engine* e, int storage) {
    PMXCHK(check_engine(e));
    if (storage != PMX_STORE_F32 && storage != PMX_STORE_F16)
        return fail(e->ctx, PMX_ERR_ARG, "unknown storage %d", storage);
    if (storage == PMX_STORE_F16 && !e->src16)
        PMXCHK(dev_alloc(e->ctx, (void**)&e->src16,
                         sizeof(uint16_t) * (size_t)e->src_layer * e->V));
    if (storage == PMX_STORE_F16 && e->src2) {
        /* the FP32 pairs are rebuilt if FP32 storage is chosen again */
        HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
        dev_free(e->src2);
        e->src2 = nullptr;
    }
    if (storage != e->storage) e->prepared = false;
    e->storage = storage;
    return PMX_OK;
}

int pmx_engine_get_storage(pmx_engine* e, int* storage) {
    PMXCHK(check_engine(e));
    if (!storage) return fail(e->ctx, PMX_ERR_ARG, "NULL argument");
    *storage = e->storage;
    return PMX_OK;
}

int pmx_engine_set_proposal_seeding(pmx_engine* e, int seed_depth) {
    PMXCHK(check_engine(e));
    e->prop_seeds = seed_depth != 0;
    return PMX_OK;
}

int pmx_engine_set_proposal_sweeps(pmx_engine* e, int sweeps) {
    PMXCHK(check_engine(e));
    if (sweeps < 0) return fail(e->ctx, PMX_ERR_ARG, "proposal sweeps must be >= 0 (got %d)", sweeps);
    e->prop_sweeps = sweeps;
    return PMX_OK;
}

int pmx_engine_set_num_iterations(pmx_engine* e, int iterations) {
    PMXCHK(check_engine(e));
    if (iterations < 1) return fail(e->ctx, PMX_ERR_ARG, "iterations must be at least 1");
    e->p.num_iterations = iterations;
    return PMX_OK;
}

int pmx_engine_cost_volume(pmx_engine* e, int view0, int view1, void* x_dev) {
    PMXCHK(check_engine(e));
    PMXCHK(require_prepared(e));
    if (!x_dev) return fail(e->ctx, PMX_ERR_ARG, "x_dev is NULL");
    if (view0 < 0 || view0 >= e->V || view1 < 0 || view1 >= e->V)
        return fail(e->ctx, PMX_ERR_ARG, "views %d, %d out of range (%d source views)", view0,
                    view1, e->V);
    PmxSweepArgs a = sweep_args(e, 0);
    HIPCHK(e->ctx, pmx_launch_cost_volume(e->ctx->stream, a, view0, view1, (uint16_t*)x_dev,
                                          (e->W + 3) / 4, (e->H + 3) / 4));
    return PMX_OK;
}

int pmx_engine_set_proposal(pmx_engine* e, const float* depth_dev, const float* spread_dev,
                            int pitch_bytes) {
    PMXCHK(check_engine(e));
    pmx_ctx* ctx = e->ctx;
    if (!depth_dev) {
        e->has_prop = false;
        return PMX_OK;
    }
    if (pitch_bytes < e->W * 4 || pitch_bytes % 4)
        return fail(ctx, PMX_ERR_ARG, "pitch %d too small or not a multiple of 4", pitch_bytes);
    hipStream_t s = ctx->stream;
    if (!e->prop_rm) {
        e->prop_rm_ld = round_up(e->W, 64);
        e->prop_cm_ld = round_up(e->H, 64);
        PMXCHK(dev_alloc(ctx, (void**)&e->prop_rm, sizeof(float) * (size_t)e->prop_rm_ld * e->H));
        PMXCHK(dev_alloc(ctx, (void**)&e->prop_cm, sizeof(float) * (size_t)e->prop_cm_ld * e->W));
        PMXCHK(dev_alloc(ctx, (void**)&e->spread_rm, sizeof(float) * (size_t)e->prop_rm_ld * e->H));
        PMXCHK(dev_alloc(ctx, (void**)&e->spread_cm, sizeof(float) * (size_t)e->prop_cm_ld * e->W));
    }
    HIPCHK(ctx, hipMemcpy2DAsync(e->prop_rm, sizeof(float) * e->prop_rm_ld, depth_dev, pitch_bytes,
                                 sizeof(float) * e->W, e->H, hipMemcpyDeviceToDevice, s));
    if (spread_dev)
        HIPCHK(ctx, hipMemcpy2DAsync(e->spread_rm, sizeof(float) * e->prop_rm_ld, spread_dev,
                                     pitch_bytes, sizeof(float) * e->W, e->H,
                                     hipMemcpyDeviceToDevice, s));
    else
        HIPCHK(ctx, hipMemsetAsync(e->spread_rm, 0, sizeof(float) * (size_t)e->prop_rm_ld * e->H, s));
    HIPCHK(ctx, pmx_launch_transpose_rec(s, e->prop_rm, e->prop_cm, e->W, e->H, 1, e->prop_rm_ld,
                                         e->prop_cm_ld));
    HIPCHK(ctx, pmx_launch_transpose_rec(s, e->spread_rm, e->spread_cm, e->W, e->H, 1,
                                         e->prop_rm_ld, e->prop_cm_ld));
    e->has_prop = true;
    if (e->prepared && e->prop_seeds) {
        if (e->layout_cm)
            HIPCHK(ctx, pmx_launch_seed_depth(s, e->prop_cm, e->prop_cm_ld, e->depth_cm,
                                              e->depth_cm_ld, e->H, e->W));
        else
            HIPCHK(ctx, pmx_launch_seed_depth(s, e->prop_rm, e->prop_rm_ld, e->depth_rm,
                                              e->depth_rm_ld, e->W, e->H));
    }
    return PMX_OK;
}

int pmx_engine_propose(pmx_engine* e, pmx_net* net, int view0, int view1, float tau) {
    PMXCHK(check_engine(e));
    PMXCHK(require_prepared(e));
    if (!net) return fail(e->ctx, PMX_ERR_ARG, "net is NULL");
    /* the cost volume runs on the engine's stream and the network on its
     * own context's: one stream orders them only when both are the same */
    if (pmx_net_context(net) != e->ctx)
        return fail(e->ctx, PMX_ERR_ARG, "the network was created on another context");
    pmx_ctx* ctx = e->ctx;
    const int W0 = (e->W + 3) / 4, H0 = (e->H + 3) / 4;
    if (view0 < 0 || view1 < 0) volume_views(e, &view0, &view1);
    if (!e->prop_x) PMXCHK(dev_alloc(ctx, &e->prop_x, sizeof(uint16_t) * (size_t)W0 * H0 * 64));
    if (!e->prop_d)
        PMXCHK(dev_alloc(ctx, (void**)&e->prop_d, 2 * sizeof(float) * (size_t)e->W * e->H));
    void* x = e->prop_x;
    float* d = e->prop_d;
    float* sp = d + (size_t)e->W * e->H;
    int rc = PMX_OK;
    if (rc == PMX_OK) rc = pmx_engine_cost_volume(e, view0, view1, x);
    if (rc == PMX_OK) rc = pmx_net_forward(net, x, 1, H0, W0);
    if (rc == PMX_OK)
        rc = pmx_net_propose(net, e->W, e->H, tau, d, e->W * 4, sp, e->W * 4, nullptr, 0);
    if (rc == PMX_OK) rc = pmx_engine_set_proposal(e, d, sp, e->W * 4);
    hipError_t es = hipStreamSynchronize(ctx->stream);
    if (rc == PMX_OK && es != hipSuccess)
        rc = fail(ctx, PMX_ERR_HIP, "%s", hipGetErrorString(es));
    return rc;
}

int pmx_engine_cost_all_views(pmx_engine* e, int rotated) {
    PMXCHK(check_engine(e));
    PMXCHK(require_prepared(e));
    Timer tm{e, {}};
    int rc = do_cost_all(e, rotated, &tm);
    tm.flush();
    return rc;
}

int pmx_engine_sweep(pmx_engine* e, int down, int rotated, int S, int record_ids) {
    PMXCHK(check_engine(e));
    PMXCHK(require_prepared(e));
    if (S < 1 || S > PMX_MAX_SAMPLES) return fail(e->ctx, PMX_ERR_ARG, "invalid num_samples %d", S);
    Timer tm{e, {}};
    int rc = do_sweep(e, down ? 1 : 0, rotated ? 1 : 0, S, record_ids != 0, &tm);
    tm.flush();
    return rc;
}

int pmx_engine_get_depth(pmx_engine* e, float* host, int pitch) {
    PMXCHK(check_engine(e));
    if (!host || pitch < e->W * 4) return fail(e->ctx, PMX_ERR_ARG, "invalid host buffer");
    PMXCHK(to_layout(e, 0, nullptr));
    HIPCHK(e->ctx, hipMemcpy2DAsync(host, pitch, e->depth_rm, (size_t)e->depth_rm_ld * 4,
                                    (size_t)e->W * 4, e->H, hipMemcpyDeviceToHost, e->ctx->stream));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return PMX_OK;
}

int pmx_engine_set_depth(pmx_engine* e, const float* host, int pitch) {
    PMXCHK(check_engine(e));
    if (!host || pitch < e->W * 4) return fail(e->ctx, PMX_ERR_ARG, "invalid host buffer");
    PMXCHK(to_layout(e, 0, nullptr));
    HIPCHK(e->ctx, hipMemcpy2DAsync(e->depth_rm, (size_t)e->depth_rm_ld * 4, host, pitch,
                                    (size_t)e->W * 4, e->H, hipMemcpyHostToDevice, e->ctx->stream));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return PMX_OK;
}

int pmx_engine_get_costs(pmx_engine* e, float* host) {
    PMXCHK(check_engine(e));
    if (!host) return fail(e->ctx, PMX_ERR_ARG, "invalid host buffer");
    PMXCHK(to_layout(e, 0, nullptr));
    const int ld = e->out_cost_pitch / 4;
    HIPCHK(e->ctx, pmx_launch_gather_costs(e->ctx->stream, e->cost_rm, e->cost_rm_ld, e->out_cost,
                                           ld, e->W, e->H, e->V));
    HIPCHK(e->ctx, hipMemcpy2DAsync(host, (size_t)e->W * 4, e->out_cost, e->out_cost_pitch,
                                    (size_t)e->W * 4, (size_t)e->H * e->V, hipMemcpyDeviceToHost,
                                    e->ctx->stream));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return PMX_OK;
}

int pmx_engine_set_costs(pmx_engine* e, const float* host) {
    PMXCHK(check_engine(e));
    if (!host) return fail(e->ctx, PMX_ERR_ARG, "invalid host buffer");
    PMXCHK(to_layout(e, 0, nullptr));
    const int ld = e->out_cost_pitch / 4;
    HIPCHK(e->ctx, hipMemcpy2DAsync(e->out_cost, e->out_cost_pitch, host, (size_t)e->W * 4,
                                    (size_t)e->W * 4, (size_t)e->H * e->V, hipMemcpyHostToDevice,
                                    e->ctx->stream));
    HIPCHK(e->ctx, pmx_launch_scatter_costs(e->ctx->stream, e->out_cost, ld, e->cost_rm,
                                            e->cost_rm_ld, e->W, e->H, e->V));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return PMX_OK;
}

int pmx_engine_get_rng(pmx_engine* e, pmx_rng_state* host) {
    PMXCHK(check_engine(e));
    if (!host) return fail(e->ctx, PMX_ERR_ARG, "invalid host buffer");
    HIPCHK(e->ctx, hipMemcpyAsync(host, e->rng, (size_t)e->rng_pitch * e->H, hipMemcpyDeviceToHost,
                                  e->ctx->stream));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return PMX_OK;
}

int pmx_engine_set_rng(pmx_engine* e, const pmx_rng_state* host) {
    PMXCHK(check_engine(e));
    if (!host) return fail(e->ctx, PMX_ERR_ARG, "invalid host buffer");
    HIPCHK(e->ctx, hipMemcpyAsync(e->rng, host, (size_t)e->rng_pitch * e->H, hipMemcpyHostToDevice,
                                  e->ctx->stream));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return PMX_OK;
}

int pmx_engine_get_used_ids(pmx_engine* e, uint8_t* host) {
    PMXCHK(check_engine(e));
    if (!host) return fail(e->ctx, PMX_ERR_ARG, "invalid host buffer");
    HIPCHK(e->ctx, hipMemcpy2DAsync(host, e->W, e->used_ids, e->ids_pitch, e->W,
                                    (size_t)e->H * e->ids_layers, hipMemcpyDeviceToHost,
                                    e->ctx->stream));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return PMX_OK;
}

int pmx_engine_get_images(pmx_engine* e, float* ref_host, float* src_host) {
    PMXCHK(check_engine(e));
    PMXCHK(require_prepared(e)); /* the pre-blurred images exist from prepare on */
    hipStream_t s = e->ctx->stream;
    if (ref_host)
        HIPCHK(e->ctx, hipMemcpy2DAsync(ref_host, (size_t)e->W * 4, e->ref, (size_t)e->ref_pitch * 4,
                                        (size_t)e->W * 4, e->H, hipMemcpyDeviceToHost, s));
    if (src_host)
        for (int v = 0; v < e->V; v++)
            HIPCHK(e->ctx,
                   hipMemcpy2DAsync(src_host + (size_t)v * e->maxW * e->maxH, (size_t)e->maxW * 4,
                                    e->src + (size_t)v * e->src_layer +
                                        (size_t)PMX_SRC_PAD * e->src_pitch + PMX_SRC_PAD,
                                    (size_t)e->src_pitch * 4, (size_t)e->maxW * 4, e->maxH,
                                    hipMemcpyDeviceToHost, s));
    HIPCHK(e->ctx, hipStreamSynchronize(s));
    return PMX_OK;
}

int pmx_engine_get_ref_sums(pmx_engine* e, float* sum_I, float* sum_II) {
    PMXCHK(check_engine(e));
    PMXCHK(require_prepared(e));
    hipStream_t s = e->ctx->stream;
    if (sum_I)
        HIPCHK(e->ctx, hipMemcpy2DAsync(sum_I, (size_t)e->W * 4, e->sum_I, (size_t)e->ref_pitch * 4,
                                        (size_t)e->W * 4, e->H, hipMemcpyDeviceToHost, s));
    if (sum_II)
        HIPCHK(e->ctx, hipMemcpy2DAsync(sum_II, (size_t)e->W * 4, e->sum_II, (size_t)e->ref_pitch * 4,
                                        (size_t)e->W * 4, e->H, hipMemcpyDeviceToHost, s));
    HIPCHK(e->ctx, hipStreamSynchronize(s));
    return PMX_OK;
}

int pmx_engine_get_homographies(pmx_engine* e, double* host) {
    PMXCHK(check_engine(e));
    if (!host) return fail(e->ctx, PMX_ERR_ARG, "invalid host buffer");
    std::memcpy(host, e->Hhost.data(), e->Hhost.size() * sizeof(double));
    return PMX_OK;
}

int pmx_engine_get_dims(pmx_engine* e, int* width, int* height, int* num_src, int* max_width,
                        int* max_height) {
    PMXCHK(check_engine(e));
    if (width) *width = e->W;
    if (height) *height = e->H;
    if (num_src) *num_src = e->V;
    if (max_width) *max_width = e->maxW;
    if (max_height) *max_height = e->maxH;
    return PMX_OK;
}

int pmx_engine_export(pmx_engine* e, float** depth, int* depth_pitch, float** costs,
                      int* cost_pitch, uint8_t** used_ids, int* ids_pitch) {
    PMXCHK(check_engine(e));
    PMXCHK(to_layout(e, 0, nullptr));
    HIPCHK(e->ctx, pmx_launch_gather_costs(e->ctx->stream, e->cost_rm, e->cost_rm_ld, e->out_cost,
                                           e->out_cost_pitch / 4, e->W, e->H, e->V));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    if (depth) *depth = e->depth_rm;
    if (depth_pitch) *depth_pitch = e->depth_rm_ld * 4;
    if (costs) *costs = e->out_cost;
    if (cost_pitch) *cost_pitch = e->out_cost_pitch;
    if (used_ids) *used_ids = e->used_ids;
    if (ids_pitch) *ids_pitch = e->ids_pitch;
    return PMX_OK;
}

int pmx_engine_export_state(pmx_engine* e, float** ref, float** sum_I, float** sum_II,
                            int* ref_pitch, pmx_rng_state** rng, int* rng_pitch) {
    PMXCHK(check_engine(e));
    PMXCHK(require_prepared(e));
    if (ref) *ref = e->ref;
    if (sum_I) *sum_I = e->sum_I;
    if (sum_II) *sum_II = e->sum_II;
    if (ref_pitch) *ref_pitch = e->ref_pitch * 4;
    if (rng) *rng = e->rng;
    if (rng_pitch) *rng_pitch = e->rng_pitch;
    return PMX_OK;
}

int pmx_engine_set_timing(pmx_engine* e, int enabled) {
    PMXCHK(check_engine(e));
    e->timing = enabled != 0;
    return PMX_OK;
}

int pmx_engine_set_sweep_waves(pmx_engine* e, int waves) {
    PMXCHK(check_engine(e));
    if (waves != 0 && waves != 1 && waves != 2 && waves != 4)
        return fail(e->ctx, PMX_ERR_ARG, "sweep waves must be 0 (auto), 1, 2 or 4 (got %d)",
                    waves);
    e->sweep_waves = waves;
    return PMX_OK;
}

int pmx_engine_get_sweep_waves(pmx_engine* e, int* waves) {
    PMXCHK(check_engine(e));
    if (!waves) return fail(e->ctx, PMX_ERR_ARG, "waves is NULL");
    *waves = e->sweep_waves;
    return PMX_OK;
}

int pmx_engine_get_stats(pmx_engine* e, pmx_stats* out) {
    PMXCHK(check_engine(e));
    if (!out) return fail(e->ctx, PMX_ERR_ARG, "out is NULL");
    unsigned long long c[4] = {0, 0, 0, 0};
    HIPCHK(e->ctx, hipMemcpyAsync(c, e->counters, sizeof(c), hipMemcpyDeviceToHost, e->ctx->stream));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    *out = e->stats_acc;
    out->ncc_evals = c[0];
    out->ncc_samples = c[1];
    out->cost_all_evals = e->cost_all_evals;
    out->cost_all_samples = e->cost_all_samples;
    return PMX_OK;
}

int pmx_engine_reset_stats(pmx_engine* e) {
    PMXCHK(check_engine(e));
    HIPCHK(e->ctx, hipMemsetAsync(e->counters, 0, sizeof(unsigned long long) * 4, e->ctx->stream));
    HIPCHK(e->ctx, hipStreamSynchronize(e->ctx->stream));
    e->stats_acc = pmx_stats{};
    e->cost_all_evals = 0;
    e->cost_all_samples = 0;
    return PMX_OK;
}

} // extern "C"

/* ------------------------------------------------------------- batches */
/*
 * Many reference views through the run<W> schedule at once
 * (src/patchMatch.cu:307-407 per view).  Each of the 4 x iterations sweep
 * steps is one persistent launch over the column strips of every view
 * (pmx_sweep_batch_kernel); cost-all-views and the state transposes stay
 * per view.  The whole schedule can be captured once in a hipGraph and
 * replayed, which removes the per-launch host cost of the ~20 launches per
 * view per iteration.
 */
struct pmx_batch {
    pmx_ctx* ctx = nullptr;
    std::vector<pmx_engine*> engs;
    int waves = 1;            /* launch shape: waves per column strip (1 or 2), or 3 */
    int grid = 0;             /* persistent workgroups; 0 = every resident slot */
    bool use_graph = true;
    int steps = 0;            /* sweep steps of one run */
    int max_V = 0;
    /* device tables, steps x: n args, 8 x (n + 1) queue starts, 8 x n first columns */
    PmxSweepArgs* d_tab = nullptr;
    int* d_qstart = nullptr;
    int* d_qcol = nullptr;
    unsigned* d_queue = nullptr; /* steps x 8 */
    std::vector<long long> sig; /* configuration the tables and graph were built for */
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

void batch_drop_tables(pmx_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    dev_free(b->d_tab);
    dev_free(b->d_qstart);
    dev_free(b->d_qcol);
    dev_free(b->d_queue);
    b->d_tab = nullptr;
    b->d_qstart = b->d_qcol = nullptr;
    b->d_queue = nullptr;
    b->sig.clear();
}

std::vector<long long> batch_signature(const pmx_batch* b) {
    std::vector<long long> s{b->waves, b->grid, b->use_graph ? 1 : 0};
    for (const pmx_engine* e : b->engs) {
        s.push_back(e->p.num_iterations);
        s.push_back(e->storage);
        s.push_back(e->has_prop ? 1 : 0);
        s.push_back(e->prop_sweeps);
    }
    return s;
}

/* (down, rotated) of the four sweeps of an iteration (src/patchMatch.cu:344-391) */
constexpr int kSweepDir[4][2] = {{1, 1}, {1, 0}, {0, 1}, {0, 0}};

int batch_build_tables(pmx_batch* b) {
    pmx_ctx* ctx = b->ctx;
    const int n = (int)b->engs.size();
    const int iters = b->engs[0]->p.num_iterations;
    b->steps = 4 * iters;
    std::vector<PmxSweepArgs> tab((size_t)b->steps * n);
    std::vector<int> qstart((size_t)b->steps * 8 * (n + 1)), qcol((size_t)b->steps * 8 * n);
    for (int k = 0; k < b->steps; k++) {
        const int it = k / 4, down = kSweepDir[k % 4][0], rot = kSweepDir[k % 4][1];
        const bool ids = it == iters - 1 && k % 4 == 3;
        /* views in order of decreasing column length */
        std::vector<int> order(n);
        for (int v = 0; v < n; v++) order[v] = v;
        auto len = [&](int v) { return rot ? b->engs[v]->W : b->engs[v]->H; };
        std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return len(p) > len(q); });
        int* qs = qstart.data() + (size_t)k * 8 * (n + 1);
        int* qc = qcol.data() + (size_t)k * 8 * n;
        for (int j = 0; j < n; j++) {
            pmx_engine* e = b->engs[order[j]];
            PmxSweepArgs a = sweep_args(e, rot);
            a.down = down;
            a.S = it == 0 ? 1 : e->p.num_samples; /* src/patchMatch.cu:328-331 */
            a.ids = ids ? e->used_ids : nullptr;
            a.sweep_id = k;
            if (e->prop_sweeps > 0 && k >= e->prop_sweeps) {
                a.prop = nullptr;
                a.prop_spread = nullptr;
            }
            tab[(size_t)k * n + j] = a;
        }
        for (int x = 0; x < 8; x++) {
            int acc = 0;
            for (int j = 0; j < n; j++) {
                const int Ws = tab[(size_t)k * n + j].Ws;
                const int lo = (int)((long)Ws * x / 8), hi = (int)((long)Ws * (x + 1) / 8);
                qs[x * (n + 1) + j] = acc;
                qc[x * n + j] = lo;
                acc += hi - lo;
            }
            qs[x * (n + 1) + n] = acc;
        }
    }
    PMXCHK(dev_alloc(ctx, (void**)&b->d_tab, sizeof(PmxSweepArgs) * tab.size()));
    PMXCHK(dev_alloc(ctx, (void**)&b->d_qstart, sizeof(int) * qstart.size()));
    PMXCHK(dev_alloc(ctx, (void**)&b->d_qcol, sizeof(int) * qcol.size()));
    PMXCHK(dev_alloc(ctx, (void**)&b->d_queue, sizeof(unsigned) * 8 * b->steps));
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(b->d_tab, tab.data(), sizeof(PmxSweepArgs) * tab.size(),
                               hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(b->d_qstart, qstart.data(), sizeof(int) * qstart.size(),
                               hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(b->d_qcol, qcol.data(), sizeof(int) * qcol.size(),
                               hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return PMX_OK;
}

PmxBatchLaunch batch_launch(const pmx_batch* b, int k) {
    const int n = (int)b->engs.size();
    PmxBatchLaunch L{};
    L.tab = b->d_tab + (size_t)k * n;
    L.qstart = b->d_qstart + (size_t)k * 8 * (n + 1);
    L.qcol = b->d_qcol + (size_t)k * 8 * n;
    L.n = n;
    L.queue = b->d_queue + 8 * k;
    L.waves = b->waves;
    L.hw = b->engs[0]->p.half_window;
    L.max_V = b->max_V;
    L.h16 = b->engs[0]->storage == PMX_STORE_F16;
    L.rot = kSweepDir[k % 4][1];
    int max_S = 1, with_prop = 0;
    for (const pmx_engine* e : b->engs) {
        max_S = std::max(max_S, e->p.num_samples);
        with_prop += (e->has_prop && (e->prop_sweeps <= 0 || k < e->prop_sweeps)) ? 1 : 0;
    }
    L.prp = with_prop > 0 ? 1 : 0;
    /* the small-view kernels fix the proposal's presence for all views */
    const bool uniform_prop = with_prop == 0 || with_prop == n;
    L.sm = (b->max_V <= 16 && max_S <= 7 && uniform_prop) ? 1 : 0;
    return L;
}

/* Workgroups of the persistent launch: every resident slot unless set. */
int batch_grid(pmx_batch* b, int* grid) {
    if (b->grid > 0) {
        *grid = b->grid;
        return PMX_OK;
    }
    PmxBatchLaunch L = batch_launch(b, 0);
    int per_cu = 0, cus = 0;
    HIPCHK(b->ctx, pmx_launch_sweep_batch(nullptr, L, &per_cu));
    HIPCHK(b->ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                         b->ctx->device));
    if (per_cu < 1) return fail(b->ctx, PMX_ERR_HIP, "batched sweep kernel cannot be resident");
    *grid = per_cu * cus;
    return PMX_OK;
}

/* The run schedule of every view, enqueued on the context's stream (the
 * caller may be capturing it).  Evaluation counts are kept by
 * pmx_batch_run, which knows whether the work ran. */
int batch_enqueue(pmx_batch* b, int grid) {
    pmx_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, pmx_launch_zero_u32(s, b->d_queue, 8 * b->steps));
    for (int k = 0; k < b->steps; k++) {
        const int rot = kSweepDir[k % 4][1];
        for (pmx_engine* e : b->engs) {
            if (k == 0) {
                const unsigned long long ce = e->cost_all_evals, cs = e->cost_all_samples;
                PMXCHK(do_cost_all(e, 1, nullptr));
                e->cost_all_evals = ce;
                e->cost_all_samples = cs;
            }
            PMXCHK(to_layout(e, rot ? 0 : 1, nullptr));
        }
        PmxBatchLaunch L = batch_launch(b, k);
        L.grid = grid;
        HIPCHK(ctx, pmx_launch_sweep_batch(s, L, nullptr));
    }
    for (pmx_engine* e : b->engs) PMXCHK(to_layout(e, 0, nullptr));
    return PMX_OK;
}

int check_batch(pmx_batch* b) {
    if (!b) return fail(nullptr, PMX_ERR_ARG, "batch is NULL");
    return PMX_OK;
}

} // namespace

extern "C" {

int pmx_batch_create(pmx_ctx* ctx, pmx_engine* const* engines, int num_views, pmx_batch** out) {
    if (!ctx) return fail(nullptr, PMX_ERR_ARG, "ctx is NULL");
    if (!out || !engines) return fail(ctx, PMX_ERR_ARG, "NULL argument to pmx_batch_create");
    *out = nullptr;
    if (num_views < 1 || num_views > 4096)
        return fail(ctx, PMX_ERR_ARG, "num_views %d out of range 1..4096", num_views);
    const pmx_engine* e0 = engines[0];
    for (int v = 0; v < num_views; v++) {
        const pmx_engine* e = engines[v];
        if (!e) return fail(ctx, PMX_ERR_ARG, "engine %d is NULL", v);
        if (e->ctx != ctx) return fail(ctx, PMX_ERR_ARG, "engine %d belongs to another context", v);
        if (e->p.half_window != e0->p.half_window || e->p.num_samples != e0->p.num_samples ||
            e->p.num_iterations != e0->p.num_iterations || e->storage != e0->storage)
            return fail(ctx, PMX_ERR_ARG,
                        "engine %d differs from engine 0 in half_window, num_samples, "
                        "num_iterations or storage", v);
        for (int u = 0; u < v; u++)
            if (engines[u] == e) return fail(ctx, PMX_ERR_ARG, "engine %d appears twice", v);
    }
    pmx_batch* b = new pmx_batch();
    b->ctx = ctx;
    b->engs.assign(engines, engines + num_views);
    for (const pmx_engine* e : b->engs) b->max_V = std::max(b->max_V, e->V);
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        if (b->ev0) (void)hipEventDestroy(b->ev0);
        delete b;
        return fail(ctx, PMX_ERR_HIP, "hipEventCreate failed");
    }
    *out = b;
    return PMX_OK;
}

int pmx_batch_destroy(pmx_batch* b) {
    if (!b) return PMX_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_drop_tables(b);
    (void)hipEventDestroy(b->ev0);
    (void)hipEventDestroy(b->ev1);
    delete b;
    return PMX_OK;
}

int pmx_batch_prepare(pmx_batch* b) {
    PMXCHK(check_batch(b));
    for (pmx_engine* e : b->engs) PMXCHK(prepare_impl(e, nullptr));
    return PMX_OK;
}

int pmx_batch_run(pmx_batch* b, float* ms) {
    PMXCHK(check_batch(b));
    pmx_ctx* ctx = b->ctx;
    for (pmx_engine* e : b->engs) {
        if (e->storage != b->engs[0]->storage || e->p.num_iterations != b->engs[0]->p.num_iterations)
            return fail(ctx, PMX_ERR_ARG, "engines of the batch no longer share storage and "
                                          "num_iterations");
        if (!e->prepared || e->sweep_count != 0 || e->layout_cm != 0)
            return fail(ctx, PMX_ERR_STATE,
                        "a batch runs freshly prepared engines: call pmx_batch_prepare first");
    }
    if (batch_signature(b) != b->sig) {
        batch_drop_tables(b);
        PMXCHK(batch_build_tables(b));
        b->sig = batch_signature(b);
    }
    hipStream_t s = ctx->stream;
    int grid = 0;
    PMXCHK(batch_grid(b, &grid));
    if (b->use_graph && !b->exec) {
        hipGraph_t g = nullptr;
        HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = batch_enqueue(b, grid);
        hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc != PMX_OK || ec != hipSuccess) {
            /* nothing ran: undo the layout bookkeeping of the capture and
             * require a new prepare */
            if (g) (void)hipGraphDestroy(g);
            for (pmx_engine* e : b->engs) {
                e->layout_cm = 0;
                e->prepared = false;
            }
            if (rc != PMX_OK) return rc;
        }
        HIPCHK(ctx, ec);
        hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) b->exec = nullptr;
        HIPCHK(ctx, ei);
        /* capture recorded the layout changes without running them */
        for (pmx_engine* e : b->engs) e->layout_cm = 0;
    }
    if (ms) HIPCHK(ctx, hipEventRecord(b->ev0, s));
    if (b->use_graph) {
        HIPCHK(ctx, hipGraphLaunch(b->exec, s));
    } else {
        const int rc = batch_enqueue(b, grid);
        if (rc != PMX_OK) {
            /* part of the schedule may be enqueued: the state is undefined
             * until the next prepare */
            for (pmx_engine* e : b->engs) e->prepared = false;
            return rc;
        }
    }
    if (ms) {
        HIPCHK(ctx, hipEventRecord(b->ev1, s));
        HIPCHK(ctx, hipEventSynchronize(b->ev1));
        HIPCHK(ctx, hipEventElapsedTime(ms, b->ev0, b->ev1));
    }
    for (pmx_engine* e : b->engs) {
        e->layout_cm = 0;
        e->sweep_count = b->steps;
        e->stats_acc.sweep_launches += b->steps;
        e->cost_all_evals += (unsigned long long)e->W * e->H * e->V;
        e->cost_all_samples += window_samples(e->W, e->H, e->p.half_window) * (unsigned long long)e->V;
    }
    return PMX_OK;
}

int pmx_batch_set_graph(pmx_batch* b, int enabled) {
    PMXCHK(check_batch(b));
    b->use_graph = enabled != 0;
    return PMX_OK;
}

int pmx_batch_set_launch(pmx_batch* b, int waves, int workgroups) {
    PMXCHK(check_batch(b));
    if (waves < 1 || waves > 3)
        return fail(b->ctx, PMX_ERR_ARG, "launch shape must be 1, 2 or 3 (got %d)", waves);
    /* every one of the 8 XCD queues needs a workgroup b with b % 8 == x */
    if (workgroups != 0 && workgroups < 8)
        return fail(b->ctx, PMX_ERR_ARG, "workgroups must be 0 (auto) or at least 8 (got %d)",
                    workgroups);
    b->waves = waves;
    b->grid = workgroups;
    return PMX_OK;
}

int pmx_batch_get_launch(pmx_batch* b, int* waves, int* workgroups) {
    PMXCHK(check_batch(b));
    if (!waves || !workgroups) return fail(b->ctx, PMX_ERR_ARG, "NULL argument");
    *waves = b->waves;
    if (b->sig.empty() && !b->d_tab) {
        /* the resident-slot count needs no tables */
        int g = b->grid;
        if (g == 0) {
            PmxBatchLaunch L{};
            L.waves = b->waves;
            L.hw = b->engs[0]->p.half_window;
            L.max_V = b->max_V;
            L.h16 = b->engs[0]->storage == PMX_STORE_F16;
            int per_cu = 0, cus = 0;
            HIPCHK(b->ctx, pmx_launch_sweep_batch(nullptr, L, &per_cu));
            HIPCHK(b->ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                                 b->ctx->device));
            g = per_cu * cus;
        }
        *workgroups = g;
        return PMX_OK;
    }
    return batch_grid(b, workgroups);
}

int pmx_batch_autotune(pmx_batch* b, int* waves, int* workgroups, float* ms) {
    PMXCHK(check_batch(b));
    const bool graph = b->use_graph;
    b->use_graph = false;
    float best = 0.0f;
    int bw = 1, bg = 0;
    int rc = PMX_OK;
    for (int w = 1; w <= 3 && rc == PMX_OK; w++) {
        b->waves = w;
        b->grid = 0;
        int full = 0;
        rc = batch_grid(b, &full);
        const int cands[3] = {full, (3 * full) / 4, full / 2};
        for (int i = 0; i < 3 && rc == PMX_OK; i++) {
            if (cands[i] < 8) continue;
            b->grid = cands[i];
            float t = 0.0f;
            rc = pmx_batch_prepare(b);
            if (rc == PMX_OK) rc = pmx_batch_run(b, &t);
            if (rc == PMX_OK && (best == 0.0f || t < best)) {
                best = t;
                bw = w;
                bg = cands[i];
            }
        }
    }
    b->use_graph = graph;
    if (rc != PMX_OK) return rc;
    b->waves = bw;
    b->grid = bg;
    if (waves) *waves = bw;
    if (workgroups) *workgroups = bg;
    if (ms) *ms = best;
    return PMX_OK;
}

} // extern "C"
