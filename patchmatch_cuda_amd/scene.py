"""Seeded synthetic scenes (ctypes front end of csrc/pmx_scene.cpp)."""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from .configs import SceneSpec

_HERE = os.path.dirname(os.path.abspath(__file__))


class _Spec(ctypes.Structure):
    _fields_ = [
        ("W", ctypes.c_int), ("H", ctypes.c_int), ("V", ctypes.c_int),
        ("n_planes", ctypes.c_int), ("n_spheres", ctypes.c_int),
        ("depth_min", ctypes.c_double), ("depth_max", ctypes.c_double),
        ("focal", ctypes.c_double), ("base_min", ctypes.c_double),
        ("base_max", ctypes.c_double), ("rot_max_deg", ctypes.c_double),
        ("noise_sigma", ctypes.c_double),
        ("scene_seed", ctypes.c_uint32), ("noise_seed", ctypes.c_uint32),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libpmx_scene.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        lib.pmx_scene_render.restype = ctypes.c_int
        lib.pmx_scene_render.argtypes = [ctypes.POINTER(_Spec), ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]
        _lib = lib
    return _lib


@dataclass
class Scene:
    spec: SceneSpec
    ref: np.ndarray        # H x W float32
    srcs: np.ndarray       # V x H x W float32
    cams: np.ndarray       # (V+1) x 21 float64: K(9) R(9) C(3); row 0 = reference
    gt_depth: np.ndarray   # H x W float32 (z-depth of the reference view)

    @property
    def K(self):
        return self.cams[:, :9].reshape(-1, 3, 3)

    @property
    def R(self):
        return self.cams[:, 9:18].reshape(-1, 3, 3)

    @property
    def C(self):
        return self.cams[:, 18:21]


def crop_reference(sc: "Scene", x0: int, y0: int, w: int, h: int):
    """The same problem restricted to a w x h window of the reference view:
    reference pixels (x0.., y0..) keep their rays (principal point shifted),
    the source views are unchanged.  Returns (ref, srcs, cams, gt)."""
    cams = sc.cams.copy()
    K = cams[0, :9].reshape(3, 3).copy()
    K[0, 2] -= x0
    K[1, 2] -= y0
    cams[0, :9] = K.ravel()
    ref = np.ascontiguousarray(sc.ref[y0:y0 + h, x0:x0 + w])
    gt = np.ascontiguousarray(sc.gt_depth[y0:y0 + h, x0:x0 + w])
    return ref, sc.srcs, cams, gt


def render(spec: SceneSpec) -> Scene:
    lib = _load()
    s = _Spec(spec.W, spec.H, spec.V, spec.n_planes, spec.n_spheres,
              spec.depth_min, spec.depth_max, spec.focal, spec.base_min,
              spec.base_max, spec.rot_max_deg, spec.noise_sigma,
              spec.scene_seed, spec.noise_seed)
    imgs = np.zeros((spec.V + 1, spec.H, spec.W), np.float32)
    gt = np.zeros((spec.H, spec.W), np.float32)
    cams = np.zeros((spec.V + 1, 21), np.float64)
    rc = lib.pmx_scene_render(ctypes.byref(s), imgs.ctypes.data, gt.ctypes.data, cams.ctypes.data)
    if rc != 0:
        raise ValueError(f"pmx_scene_render failed ({rc})")
    return Scene(spec, imgs[0].copy(), imgs[1:].copy(), cams, gt)
