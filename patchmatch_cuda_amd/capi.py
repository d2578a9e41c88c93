"""ctypes binding of the C ABI in include/pmx.h (libpmx.so).

The product path is libpmx.so; there is no Python or CPU fallback: if the
library is missing or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PMX_LIBRARY selects another build of the same C ABI (e.g. an instrumented one).
LIB_PATH = os.environ.get("PMX_LIBRARY") or os.path.join(_HERE, "libpmx.so")

PMX_OK = 0
ERR_NAMES = {1: "PMX_ERR_ARG", 2: "PMX_ERR_HIP", 3: "PMX_ERR_ALLOC", 4: "PMX_ERR_STATE",
             5: "PMX_ERR_NO_DEVICE"}
IMG_F32, IMG_U8 = 0, 1
STORE_F32, STORE_F16 = 0, 1
MEMCPY_H2D, MEMCPY_D2H, MEMCPY_D2D = 1, 2, 3

RNG_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4"), ("z", "<u4"), ("w", "<u4"),
                      ("v", "<u4"), ("d", "<u4")])


class Camera(ctypes.Structure):
    _fields_ = [("K", ctypes.c_double * 9), ("R", ctypes.c_double * 9), ("C", ctypes.c_double * 3)]

    @classmethod
    def from_row(cls, row):
        row = np.asarray(row, np.float64)
        return cls((ctypes.c_double * 9)(*row[:9]), (ctypes.c_double * 9)(*row[9:18]),
                   (ctypes.c_double * 3)(*row[18:21]))


class Params(ctypes.Structure):
    _fields_ = [
        ("half_window", ctypes.c_int), ("num_samples", ctypes.c_int),
        ("num_iterations", ctypes.c_int), ("depth_near", ctypes.c_float),
        ("depth_far", ctypes.c_float), ("spm_alpha", ctypes.c_float),
        ("gaussian_sigma", ctypes.c_float), ("blur_mode", ctypes.c_int),
        ("init_mode", ctypes.c_int), ("orientation_x", ctypes.c_float),
        ("orientation_z", ctypes.c_float), ("block_x", ctypes.c_int),
        ("block_y", ctypes.c_int), ("rng_seed", ctypes.c_uint32),
    ]

    @classmethod
    def from_run_params(cls, rp):
        return cls(rp.half_window, rp.num_samples, rp.num_iterations, rp.depth_near,
                   rp.depth_far, rp.spm_alpha, rp.gaussian_sigma, rp.blur_mode, rp.init_mode,
                   rp.orientation_x, rp.orientation_z, rp.block_x, rp.block_y, rp.rng_seed)


class Stats(ctypes.Structure):
    _fields_ = [
        ("ncc_evals", ctypes.c_uint64), ("ncc_samples", ctypes.c_uint64),
        ("cost_all_evals", ctypes.c_uint64), ("cost_all_samples", ctypes.c_uint64),
        ("sweep_launches", ctypes.c_int), ("transpose_launches", ctypes.c_int),
        ("ms_sweep", ctypes.c_double), ("ms_cost_all", ctypes.c_double),
        ("ms_transpose", ctypes.c_double), ("ms_prepare", ctypes.c_double),
        ("ms_sweep_max", ctypes.c_double),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float

# name -> argtypes (all return int unless listed in _RESTYPES)
SIGNATURES = {
    "pmx_ctx_create": [_i, ctypes.POINTER(_vp)],
    "pmx_ctx_destroy": [_vp],
    "pmx_last_error": [_vp],
    "pmx_ctx_sync": [_vp],
    "pmx_ctx_stream": [_vp],
    "pmx_device_count": [ctypes.POINTER(_i)],
    "pmx_malloc_pitch": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_size_t),
                         ctypes.c_size_t, ctypes.c_size_t],
    "pmx_malloc": [_vp, ctypes.POINTER(_vp), ctypes.c_size_t],
    "pmx_free": [_vp, _vp],
    "pmx_memcpy2d": [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, ctypes.c_size_t,
                     ctypes.c_size_t, _i],
    "pmx_memset": [_vp, _vp, _i, ctypes.c_size_t],
    "pmx_transpose_f32": [_vp, _vp, _vp, _i, _i, _i, _i],
    "pmx_rng_init": [_vp, _vp, _i, _i, _i, _i, _i, ctypes.c_uint32],
    "pmx_rng_fill_depth": [_vp, _vp, _i, _vp, _i, _i, _i, _f, _f, _i],
    "pmx_box_sums": [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _i],
    "pmx_blur": [_vp, _vp, _i, _i, _i, _i, _f, _i],
    "pmx_blur_noalloc": [_vp, _vp, _i, _i, _i, _i, _f, _i],
    "pmx_compose_homographies": [_vp, ctypes.POINTER(Camera), ctypes.POINTER(Camera), _i, _f, _f,
                                 _vp],
    "pmx_validate_params": [ctypes.POINTER(Params), _i],
    "pmx_engine_create": [_vp, ctypes.POINTER(Params), _i, _i, _i, _vp, _i, _i,
                          ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_vp),
                          ctypes.POINTER(_i), ctypes.POINTER(Camera), ctypes.POINTER(Camera),
                          ctypes.POINTER(_vp)],
    "pmx_engine_destroy": [_vp],
    "pmx_engine_prepare": [_vp],
    "pmx_engine_run": [_vp],
    "pmx_engine_cost_all_views": [_vp, _i],
    "pmx_engine_sweep": [_vp, _i, _i, _i, _i],
    "pmx_engine_get_depth": [_vp, _vp, _i],
    "pmx_engine_set_depth": [_vp, _vp, _i],
    "pmx_engine_get_costs": [_vp, _vp],
    "pmx_engine_set_costs": [_vp, _vp],
    "pmx_engine_get_rng": [_vp, _vp],
    "pmx_engine_set_rng": [_vp, _vp],
    "pmx_engine_get_used_ids": [_vp, _vp],
    "pmx_engine_get_images": [_vp, _vp, _vp],
    "pmx_engine_get_ref_sums": [_vp, _vp, _vp],
    "pmx_engine_get_homographies": [_vp, _vp],
    "pmx_engine_get_dims": [_vp] + [ctypes.POINTER(_i)] * 5,
    "pmx_engine_export": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_i), ctypes.POINTER(_vp),
                          ctypes.POINTER(_i), ctypes.POINTER(_vp), ctypes.POINTER(_i)],
    "pmx_engine_export_state": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                ctypes.POINTER(_vp), ctypes.POINTER(_i), ctypes.POINTER(_vp),
                                ctypes.POINTER(_i)],
    "pmx_engine_set_timing": [_vp, _i],
    "pmx_engine_set_sweep_waves": [_vp, _i],
    "pmx_engine_get_sweep_waves": [_vp, ctypes.POINTER(_i)],
    "pmx_engine_get_stats": [_vp, ctypes.POINTER(Stats)],
    "pmx_engine_reset_stats": [_vp],
    # learned proposal
    "pmx_engine_cost_volume": [_vp, _i, _i, _vp],
    "pmx_engine_set_proposal": [_vp, _vp, _vp, _i],
    "pmx_engine_volume_views": [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "pmx_engine_set_num_iterations": [_vp, _i],
    "pmx_engine_set_proposal_seeding": [_vp, _i],
    "pmx_engine_set_proposal_sweeps": [_vp, _i],
    "pmx_engine_propose": [_vp, _vp, _i, _i, _f],
    "pmx_engine_set_storage": [_vp, _i],
    "pmx_engine_footprint": [ctypes.POINTER(Params), _i, _i, _i, _vp, _vp, _i, _i,
                             ctypes.POINTER(ctypes.c_size_t)],
    "pmx_batch_create": [_vp, _vp, _i, ctypes.POINTER(_vp)],
    "pmx_batch_destroy": [_vp],
    "pmx_batch_prepare": [_vp],
    "pmx_batch_run": [_vp, ctypes.POINTER(_f)],
    "pmx_batch_set_graph": [_vp, _i],
    "pmx_batch_set_launch": [_vp, _i, _i],
    "pmx_batch_get_launch": [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "pmx_batch_autotune": [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_f)],
    "pmx_engine_get_storage": [_vp, ctypes.POINTER(_i)],
    "pmx_net_create": [_vp, _i, _i, _i, _f, _f, ctypes.c_ulonglong, ctypes.POINTER(_vp)],
    "pmx_net_destroy": [_vp],
    "pmx_net_num_params": [_vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
    "pmx_net_layer_offsets": [_vp, _vp, _vp],
    "pmx_net_get_params": [_vp, _vp, ctypes.c_size_t],
    "pmx_net_set_params": [_vp, _vp, ctypes.c_size_t],
    "pmx_net_get_grads": [_vp, _vp, ctypes.c_size_t],
    "pmx_net_zero_grads": [_vp],
    "pmx_net_forward": [_vp, _vp, _i, _i, _i],
    "pmx_net_get_head": [_vp, _vp, ctypes.c_size_t],
    "pmx_net_backward": [_vp, _vp, _vp, _f, _vp],
    "pmx_net_adam": [_vp, _f, _f, _f, _f],
    "pmx_net_train_step": [_vp, _vp, _vp, _i, _i, _i, _f, _f, _vp],
    "pmx_net_propose": [_vp, _i, _i, _f, _vp, _i, _vp, _i, _vp, _i],
    "pmx_cnn_conv_fwd": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i],
    "pmx_cnn_conv_bwd_data": [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i],
    "pmx_cnn_conv_bwd_weight": [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i],
    "pmx_cnn_weight_t": [_vp, _vp, _vp, _i, _i],
    "pmx_cnn_target": [_vp, _vp, _i, _i, _i, _vp],
}
_RESTYPES = {"pmx_last_error": ctypes.c_char_p, "pmx_ctx_stream": _vp}

_lib = None


class PmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: the HIP engine is not built "
                               "(run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def last_error(ctx=None) -> str:
    m = lib().pmx_last_error(ctx)
    return m.decode() if m else ""


def check(rc, ctx=None):
    if rc != PMX_OK:
        raise PmxError(rc, last_error(ctx))


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().pmx_device_count(ctypes.byref(n)))
    return n.value


class Context:
    def __init__(self, device=0):
        h = _vp()
        check(lib().pmx_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self._deps = weakref.WeakSet()

    def track(self, obj):
        """Registers an object that holds resources of this context (engine,
        batch, network, device array): close() releases it first, so nothing
        outlives the context it points into."""
        self._deps.add(obj)

    def close(self):
        if getattr(self, "h", None):
            # batches borrow engines; arrays and engines go before the context
            for dep in sorted(list(getattr(self, "_deps", ())), key=lambda d: d.CLOSE_ORDER):
                dep.close()
            lib().pmx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def check(self, rc):
        check(rc, self.h)

    def sync(self):
        self.check(lib().pmx_ctx_sync(self.h))

    @property
    def stream(self):
        return lib().pmx_ctx_stream(self.h)


class DeviceArray:
    """Device memory from the engine's own allocator (pmx_malloc), with
    synchronous uploads/downloads and stream-ordered device copies on the
    context's stream.  dtype is the host view: float32, uint16 (bf16 bits),
    uint8."""
    CLOSE_ORDER = 2

    def __init__(self, ctx: "Context", shape, dtype=np.float32, zero=True):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape))
        self.nbytes = self.size * self.dtype.itemsize
        h = _vp()
        ctx.check(lib().pmx_malloc(ctx.h, ctypes.byref(h), max(self.nbytes, 1)))
        self.h = h
        ctx.track(self)
        if zero:
            ctx.check(lib().pmx_memset(ctx.h, self.h, 0, self.nbytes))

    @property
    def ptr(self) -> int:
        return self.h.value

    def at(self, index) -> int:
        """Device address of element `index` along the first axis."""
        return self.h.value + int(index) * (self.nbytes // self.shape[0])

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.size, (a.shape, self.shape)
        self.ctx.check(lib().pmx_memcpy2d(self.ctx.h, self.h, self.nbytes, _vp(a.ctypes.data),
                                          self.nbytes, self.nbytes, 1, MEMCPY_H2D))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx.check(lib().pmx_memcpy2d(self.ctx.h, _vp(out.ctypes.data), self.nbytes, self.h,
                                          self.nbytes, self.nbytes, 1, MEMCPY_D2H))
        return out

    def copy_rows(self, dst_index, src, src_index, count=1):
        """Stream-ordered copy of `count` leading-axis entries from src."""
        row = self.nbytes // self.shape[0]
        assert row == src.nbytes // src.shape[0]
        assert 0 <= dst_index and dst_index + count <= self.shape[0]
        assert 0 <= src_index and src_index + count <= src.shape[0]
        n = row * count
        self.ctx.check(lib().pmx_memcpy2d(self.ctx.h, _vp(self.at(dst_index)), n,
                                          _vp(src.at(src_index)), n, n, 1, MEMCPY_D2D))

    def close(self):
        if getattr(self, "h", None):
            lib().pmx_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        self.close()


def to_bf16(a) -> np.ndarray:
    """float -> bf16 bits (uint16), round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16(b) -> np.ndarray:
    return (np.ascontiguousarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class Engine:
    """One reference view against V source views, resident on the GPU."""
    CLOSE_ORDER = 1

    def __init__(self, ctx: Context, params, ref, srcs, cams, image_format=IMG_F32):
        self.ctx = ctx
        ref = np.ascontiguousarray(ref)
        srcs = [np.ascontiguousarray(s) for s in srcs]
        dt = np.float32 if image_format == IMG_F32 else np.uint8
        ref = ref.astype(dt, copy=False)
        srcs = [s.astype(dt, copy=False) for s in srcs]
        self.H, self.W = ref.shape
        self.V = len(srcs)
        self.params = params if isinstance(params, Params) else Params.from_run_params(params)
        esz = ref.itemsize
        ws = (ctypes.c_int * self.V)(*[s.shape[1] for s in srcs])
        hs = (ctypes.c_int * self.V)(*[s.shape[0] for s in srcs])
        ps = (ctypes.c_int * self.V)(*[s.shape[1] * esz for s in srcs])
        ptrs = (ctypes.c_void_p * self.V)(*[s.ctypes.data for s in srcs])
        cams = np.asarray(cams, np.float64)
        rc_cam = Camera.from_row(cams[0])
        sc = (Camera * self.V)(*[Camera.from_row(cams[v + 1]) for v in range(self.V)])
        h = _vp()
        self._keep = (ref, srcs)
        ctx.check(lib().pmx_engine_create(ctx.h, ctypes.byref(self.params), image_format, self.W,
                                          self.H, _p(ref), self.W * esz, self.V, ws, hs, ptrs, ps,
                                          ctypes.byref(rc_cam), sc, ctypes.byref(h)))
        self.h = h
        d = [ctypes.c_int() for _ in range(5)]
        ctx.check(lib().pmx_engine_get_dims(self.h, *[ctypes.byref(x) for x in d]))
        self.maxW, self.maxH = d[3].value, d[4].value
        self.S = self.params.num_samples
        ctx.track(self)

    def close(self):
        if getattr(self, "h", None):
            lib().pmx_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _c(self, rc):
        self.ctx.check(rc)

    def prepare(self):
        self._c(lib().pmx_engine_prepare(self.h))

    def run(self):
        self._c(lib().pmx_engine_run(self.h))

    def cost_all_views(self, rotated=True):
        self._c(lib().pmx_engine_cost_all_views(self.h, int(rotated)))

    def sweep(self, down, rotated, S, record_ids=False):
        self._c(lib().pmx_engine_sweep(self.h, int(down), int(rotated), int(S), int(record_ids)))

    def depth(self):
        out = np.zeros((self.H, self.W), np.float32)
        self._c(lib().pmx_engine_get_depth(self.h, _p(out), self.W * 4))
        return out

    def set_depth(self, d):
        d = np.ascontiguousarray(d, np.float32)
        self._c(lib().pmx_engine_set_depth(self.h, _p(d), self.W * 4))

    def costs(self):
        out = np.zeros((self.V, self.H, self.W), np.float32)
        self._c(lib().pmx_engine_get_costs(self.h, _p(out)))
        return out

    def set_costs(self, c):
        c = np.ascontiguousarray(c, np.float32)
        assert c.shape == (self.V, self.H, self.W)
        self._c(lib().pmx_engine_set_costs(self.h, _p(c)))

    def rng(self):
        out = np.zeros((self.H, self.W), RNG_DTYPE)
        self._c(lib().pmx_engine_get_rng(self.h, _p(out)))
        return out

    def set_rng(self, st):
        st = np.ascontiguousarray(st)
        assert st.dtype == RNG_DTYPE and st.shape == (self.H, self.W)
        self._c(lib().pmx_engine_set_rng(self.h, _p(st)))

    def used_ids(self):
        out = np.zeros((self.S, self.H, self.W), np.uint8)
        self._c(lib().pmx_engine_get_used_ids(self.h, _p(out)))
        return out

    def images(self):
        ref = np.zeros((self.H, self.W), np.float32)
        srcs = np.zeros((self.V, self.maxH, self.maxW), np.float32)
        self._c(lib().pmx_engine_get_images(self.h, _p(ref), _p(srcs)))
        return ref, srcs

    def ref_sums(self):
        a = np.zeros((self.H, self.W), np.float32)
        b = np.zeros((self.H, self.W), np.float32)
        self._c(lib().pmx_engine_get_ref_sums(self.h, _p(a), _p(b)))
        return a, b

    def homographies(self):
        out = np.zeros((self.V, 18), np.float64)
        self._c(lib().pmx_engine_get_homographies(self.h, _p(out)))
        return out

    def set_timing(self, on=True):
        self._c(lib().pmx_engine_set_timing(self.h, int(on)))

    def set_sweep_waves(self, n):
        self._c(lib().pmx_engine_set_sweep_waves(self.h, int(n)))

    def sweep_waves(self):
        n = ctypes.c_int()
        self._c(lib().pmx_engine_get_sweep_waves(self.h, ctypes.byref(n)))
        return n.value

    def stats(self):
        s = Stats()
        self._c(lib().pmx_engine_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def cost_volume(self, view0, view1, x_ptr):
        """Plane-sweep cost volume (network input) into device bf16 memory at x_ptr."""
        self._c(lib().pmx_engine_cost_volume(self.h, view0, view1, _vp(x_ptr)))

    def set_proposal(self, depth_ptr, pitch_bytes, spread_ptr=None):
        """Per-pixel proposal depth (device float plane, 0 = none) and optional
        spread in inverse depth; depth_ptr None clears it."""
        self._c(lib().pmx_engine_set_proposal(self.h, _vp(depth_ptr) if depth_ptr else None,
                                              _vp(spread_ptr) if spread_ptr else None,
                                              pitch_bytes))

    def volume_views(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._c(lib().pmx_engine_volume_views(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_proposal_seeding(self, on=True):
        self._c(lib().pmx_engine_set_proposal_seeding(self.h, int(bool(on))))

    def set_proposal_sweeps(self, n):
        """The proposal joins only the first n sweeps of a run (0: all)."""
        self._c(lib().pmx_engine_set_proposal_sweeps(self.h, int(n)))

    def set_num_iterations(self, n):
        self._c(lib().pmx_engine_set_num_iterations(self.h, n))

    def propose(self, net, view0=-1, view1=-1, tau=0.0):
        """Cost volume -> network -> proposal, installed for the next run."""
        self._c(lib().pmx_engine_propose(self.h, net.h, view0, view1, tau))

    def reset_stats(self):
        self._c(lib().pmx_engine_reset_stats(self.h))

    def set_storage(self, storage):
        """STORE_F32 or STORE_F16 texels for the NCC; needs a new prepare()."""
        self._c(lib().pmx_engine_set_storage(self.h, int(storage)))

    def storage(self):
        v = ctypes.c_int()
        self._c(lib().pmx_engine_get_storage(self.h, ctypes.byref(v)))
        return v.value


def engine_footprint(params, width, height, src_sizes, storage=STORE_F32, proposal=False) -> int:
    """Device bytes of an engine (pmx_engine_footprint); src_sizes: [(w, h)]."""
    ws = np.ascontiguousarray([w for w, _ in src_sizes], np.int32)
    hs = np.ascontiguousarray([h for _, h in src_sizes], np.int32)
    p = Params.from_run_params(params)
    out = ctypes.c_size_t()
    check(lib().pmx_engine_footprint(ctypes.byref(p), int(width), int(height), len(src_sizes),
                                     _vp(ws.ctypes.data), _vp(hs.ctypes.data), int(storage),
                                     int(bool(proposal)), ctypes.byref(out)))
    return out.value


class Batch:
    """Several engines run together (pmx_batch_*): one persistent sweep
    launch per sweep step over every view, replayed from a hipGraph."""
    CLOSE_ORDER = 0

    def __init__(self, ctx: Context, engines):
        self.ctx = ctx
        self.engines = list(engines)
        arr = (_vp * len(self.engines))(*[e.h for e in self.engines])
        h = _vp()
        ctx.check(lib().pmx_batch_create(ctx.h, arr, len(self.engines), ctypes.byref(h)))
        self.h = h
        ctx.track(self)

    def close(self):
        if getattr(self, "h", None):
            lib().pmx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def prepare(self):
        self.ctx.check(lib().pmx_batch_prepare(self.h))

    def run(self, timed=False):
        """Runs the schedule; returns its duration in ms when timed."""
        ms = _f()
        self.ctx.check(lib().pmx_batch_run(self.h, ctypes.byref(ms) if timed else None))
        return ms.value if timed else None

    def set_graph(self, on=True):
        self.ctx.check(lib().pmx_batch_set_graph(self.h, int(bool(on))))

    def set_launch(self, waves, workgroups=0):
        self.ctx.check(lib().pmx_batch_set_launch(self.h, int(waves), int(workgroups)))

    def launch(self):
        w, g = _i(), _i()
        self.ctx.check(lib().pmx_batch_get_launch(self.h, ctypes.byref(w), ctypes.byref(g)))
        return w.value, g.value

    def autotune(self):
        """(waves, workgroups, ms) of the fastest launch shape, now in use."""
        w, g, ms = _i(), _i(), _f()
        self.ctx.check(lib().pmx_batch_autotune(self.h, ctypes.byref(w), ctypes.byref(g),
                                                ctypes.byref(ms)))
        return w.value, g.value, ms.value
