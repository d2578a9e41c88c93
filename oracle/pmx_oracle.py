"""ctypes front end of the FP64 CPU oracle (oracle/pmx_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py -- always as the checker or the CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpmx_oracle.so")

RNG_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4"), ("z", "<u4"), ("w", "<u4"),
                      ("v", "<u4"), ("d", "<u4")])


class _Rng(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("x", "y", "z", "w", "v", "d")]


class Params(ctypes.Structure):
    _fields_ = [
        ("W", ctypes.c_int), ("H", ctypes.c_int), ("V", ctypes.c_int),
        ("srcW", ctypes.c_int), ("srcH", ctypes.c_int), ("hw", ctypes.c_int),
        ("near_d", ctypes.c_double), ("far_d", ctypes.c_double),
        ("alpha", ctypes.c_double), ("oX", ctypes.c_double), ("oZ", ctypes.c_double),
        ("invK", ctypes.c_double * 4),
    ]


class _Trace(ctypes.Structure):
    _fields_ = [("view_margin", ctypes.c_void_p), ("cost_margin", ctypes.c_void_p),
                ("choice", ctypes.c_void_p), ("near_clamp", ctypes.c_void_p),
                ("cdf", ctypes.c_void_p), ("rand_depth", ctypes.c_void_p),
                ("cand_cost", ctypes.c_void_p), ("lpost_max", ctypes.c_void_p)]


_lib = None
_vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "pmxo_rng_seed": (None, [_vp, ctypes.c_uint32, ctypes.c_uint32]),
            "pmxo_rng_next": (ctypes.c_uint32, [_vp]),
            "pmxo_rng_uniform": (ctypes.c_double, [_vp]),
            "pmxo_seed_id": (ctypes.c_uint32, [ctypes.c_int] * 5),
            "pmxo_rng_init_grid": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_uint32]),
            "pmxo_init_depth": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, _vp]),
            "pmxo_blur_taps": (ctypes.c_int, [ctypes.c_double, ctypes.c_int, _vp]),
            "pmxo_blur": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, _vp]),
            "pmxo_box_sums": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
            "pmxo_transpose": (None, [_vp, ctypes.c_int, ctypes.c_int, _vp]),
            "pmxo_compose": (None, [_vp] * 6 + [ctypes.c_double, ctypes.c_double, _vp, _vp]),
            "pmxo_problem_create": (_vp, [ctypes.POINTER(Params), _vp, _vp, _vp]),
            "pmxo_problem_destroy": (None, [_vp]),
            "pmxo_problem_ref_sums": (None, [_vp, _vp, _vp]),
            "pmxo_ncc": (ctypes.c_double, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_int]),
            "pmxo_cost_all_views": (None, [_vp, _vp, _vp, ctypes.c_int]),
            "pmxo_sweep": (None, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _vp, _vp]),
            "pmxo_run": (None, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
            "pmxo_counters": (None, [_vp, _vp, _vp]),
            "pmxo_problem_set_proposal": (None, [_vp, _vp, _vp]),
            "pmxo_problem_set_proposal_sweeps": (None, [_vp, ctypes.c_int]),
            "pmxo_problem_set_sweep_id": (None, [_vp, ctypes.c_int]),
            "pmxo_proposal_draw": (ctypes.c_double, [ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                     ctypes.c_double, ctypes.c_int]),
            "pmxo_pick_view": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double]),
            "pmxo_hmm_posteriors": (None, [_vp, ctypes.c_int, ctypes.c_double, _vp]),
            "pmxo_set_threads": (None, [ctypes.c_int]),
            "pmxo_get_threads": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


def set_threads(n: int):
    lib().pmxo_set_threads(int(n))


def get_threads() -> int:
    return int(lib().pmxo_get_threads())


# ---------------------------------------------------------------- rng
def rng_seed(id_: int, seed: int) -> np.ndarray:
    st = np.zeros(1, RNG_DTYPE)
    lib().pmxo_rng_seed(_p(st), id_, seed)
    return st


def rng_next(st: np.ndarray) -> int:
    return int(lib().pmxo_rng_next(_p(st)))


def rng_uniform(st: np.ndarray) -> float:
    return float(lib().pmxo_rng_uniform(_p(st)))


def seed_id(x, y, width, bx=32, by=16) -> int:
    return int(lib().pmxo_seed_id(x, y, width, bx, by))


def rng_init_grid(W, H, bx=32, by=16, seed=7) -> np.ndarray:
    st = np.zeros((H, W), RNG_DTYPE)
    lib().pmxo_rng_init_grid(_p(st), W, H, bx, by, seed)
    return st


def init_depth(states: np.ndarray, near, far, mode) -> np.ndarray:
    H, W = states.shape
    d = np.zeros((H, W), np.float64)
    lib().pmxo_init_depth(_p(states), W, H, near, far, mode, _p(d))
    return d


# --------------------------------------------------------- image ops
def blur_taps(sigma, mode):
    t = np.zeros(45, np.float64)
    w = lib().pmxo_blur_taps(sigma, mode, _p(t))
    return t[:w].copy()


def blur(img, sigma, mode):
    img = np.ascontiguousarray(img, np.float64)
    H, W = img.shape
    out = np.zeros_like(img)
    lib().pmxo_blur(_p(img), W, H, sigma, mode, _p(out))
    return out


def box_sums(img, hw):
    img = np.ascontiguousarray(img, np.float64)
    H, W = img.shape
    a = np.zeros_like(img)
    b = np.zeros_like(img)
    lib().pmxo_box_sums(_p(img), W, H, hw, _p(a), _p(b))
    return a, b


def transpose(img):
    img = np.ascontiguousarray(img, np.float64)
    H, W = img.shape
    out = np.zeros((W, H), np.float64)
    lib().pmxo_transpose(_p(img), W, H, _p(out))
    return out


def compose(Kr, Rr, Cr, Ks, Rs, Cs, oX=0.0, oZ=1.0):
    arrs = [np.ascontiguousarray(a, np.float64) for a in (Kr, Rr, Cr, Ks, Rs, Cs)]
    H1 = np.zeros(9)
    H2 = np.zeros(9)
    lib().pmxo_compose(*[_p(a) for a in arrs], oX, oZ, _p(H1), _p(H2))
    return H1.reshape(3, 3), H2.reshape(3, 3)


def homographies(cams: np.ndarray, oX=0.0, oZ=1.0) -> np.ndarray:
    """V x 18 (H1 then H2) for cams[(V+1) x 21] with row 0 the reference."""
    K, R, C = cams[:, :9], cams[:, 9:18], cams[:, 18:21]
    out = np.zeros((cams.shape[0] - 1, 18))
    for v in range(1, cams.shape[0]):
        H1, H2 = compose(K[0], R[0], C[0], K[v], R[v], C[v], oX, oZ)
        out[v - 1, :9] = H1.ravel()
        out[v - 1, 9:] = H2.ravel()
    return out


def pick_view(cdf, u) -> int:
    cdf = np.ascontiguousarray(cdf, np.float64)
    return int(lib().pmxo_pick_view(_p(cdf), len(cdf), u))


def hmm_posteriors(costs, alpha):
    costs = np.ascontiguousarray(costs, np.float64)
    out = np.zeros(len(costs))
    lib().pmxo_hmm_posteriors(_p(costs), len(costs), alpha, _p(out))
    return out


def proposal_draw(depth, spread, x, y, W, near, far, sweep_id) -> float:
    return float(lib().pmxo_proposal_draw(depth, spread, x, y, W, near, far, sweep_id))


def inv_k4(K: np.ndarray):
    Ki = np.linalg.inv(np.asarray(K, np.float64).reshape(3, 3))
    return [Ki[0, 0], Ki[0, 2], Ki[1, 1], Ki[1, 2]]


class Problem:
    """One reference view against V source views (images already blurred)."""

    def __init__(self, ref, srcs, H1H2, hw, near, far, alpha, invK, oX=0.0, oZ=1.0):
        ref = np.ascontiguousarray(ref, np.float64)
        srcs = np.ascontiguousarray(srcs, np.float64)
        self.H, self.W = ref.shape
        self.V, self.srcH, self.srcW = srcs.shape
        self.hw = hw
        p = Params(self.W, self.H, self.V, self.srcW, self.srcH, hw, near, far, alpha,
                   oX, oZ, (ctypes.c_double * 4)(*invK))
        self._keep = (ref, srcs, np.ascontiguousarray(H1H2, np.float64))
        self._h = lib().pmxo_problem_create(ctypes.byref(p), _p(ref), _p(srcs), _p(self._keep[2]))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pmxo_problem_destroy(h)
            self._h = None

    def ref_sums(self):
        a = np.zeros((self.H, self.W))
        b = np.zeros((self.H, self.W))
        lib().pmxo_problem_ref_sums(self._h, _p(a), _p(b))
        return a, b

    def ncc(self, view, row_s, col_s, depth, rotated=False) -> float:
        return float(lib().pmxo_ncc(self._h, view, row_s, col_s, depth, int(rotated)))

    def frame(self, rotated):
        return (self.H, self.W) if rotated else (self.W, self.H)  # (Ws, Hs)

    def cost_all_views(self, depth_s, rotated=True):
        Ws, Hs = self.frame(rotated)
        depth_s = np.ascontiguousarray(depth_s, np.float64)
        assert depth_s.shape == (Hs, Ws)
        cost = np.zeros((self.V, Hs, Ws))
        lib().pmxo_cost_all_views(self._h, _p(depth_s), _p(cost), int(rotated))
        return cost

    def sweep(self, depth_s, cost_s, states, down, rotated, S, record_ids=False, trace=False):
        """In-place sweep over sweep-frame arrays; returns (used_ids, trace dict)."""
        Ws, Hs = self.frame(rotated)
        assert depth_s.shape == (Hs, Ws) and depth_s.dtype == np.float64 and depth_s.flags.c_contiguous
        assert cost_s.shape == (self.V, Hs, Ws) and cost_s.dtype == np.float64
        assert states.shape == (self.H, self.W) and states.dtype == RNG_DTYPE
        ids = np.zeros((S, Hs, Ws), np.uint8) if record_ids else None
        tr = None
        trs = None
        if trace:
            tr = dict(view_margin=np.zeros((Hs, Ws)), cost_margin=np.zeros((Hs, Ws)),
                      choice=np.zeros((Hs, Ws), np.uint8), near_clamp=np.zeros((Hs, Ws), np.uint8),
                      lpost_max=np.zeros((Hs, Ws)))
            if trace == "full":
                tr.update(cdf=np.zeros((Hs, Ws, self.V)), rand_depth=np.zeros((Hs, Ws)),
                          cand_cost=np.zeros((Hs, Ws, 4)))
            trs = _Trace(_p(tr["view_margin"]), _p(tr["cost_margin"]), _p(tr["choice"]),
                         _p(tr["near_clamp"]), _p(tr.get("cdf")), _p(tr.get("rand_depth")),
                         _p(tr.get("cand_cost")), _p(tr["lpost_max"]))
        lib().pmxo_sweep(self._h, _p(depth_s), _p(cost_s), _p(states), int(down), int(rotated), S,
                         _p(ids), ctypes.byref(trs) if trs is not None else None)
        return ids, tr

    def run(self, depth, states, S, iterations, record_ids=True):
        depth = np.ascontiguousarray(depth, np.float64).copy()
        cost = np.zeros((self.V, self.H, self.W))
        ids = np.zeros((S, self.H, self.W), np.uint8) if record_ids else None
        lib().pmxo_run(self._h, _p(depth), _p(cost), _p(states), S, iterations, _p(ids))
        return depth, cost, ids

    def set_proposal(self, prop, spread=None):
        """Proposal depths (H x W, original frame, 0 = none) and spreads in
        inverse depth (or None = 0); prop None removes it.  Resets the sweep
        count that keys the proposal draws."""
        if prop is None:
            lib().pmxo_problem_set_proposal(self._h, None, None)
            self._prop = None
            return
        self._prop = np.ascontiguousarray(prop, np.float64)
        assert self._prop.shape == (self.H, self.W)
        self._spread = None if spread is None else np.ascontiguousarray(spread, np.float64)
        lib().pmxo_problem_set_proposal(self._h, _p(self._prop),
                                        _p(self._spread) if self._spread is not None else None)

    def set_sweep_id(self, k):
        lib().pmxo_problem_set_sweep_id(self._h, int(k))

    def set_proposal_sweeps(self, n):
        """The proposal joins only the first n sweeps (0: all)."""
        lib().pmxo_problem_set_proposal_sweeps(self._h, int(n))

    def counters(self):
        a = ctypes.c_uint64()
        b = ctypes.c_uint64()
        lib().pmxo_counters(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value
