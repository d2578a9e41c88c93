"""Named PatchMatch workloads: scene geometry plus the run parameters.

The reference reads its parameters from an INI file (src/main.cpp:42-91:
halfWindowSize, numOfSamples, SPMAlpha, gaussianSigma, numOfIterations,
depthRangeNear/Far, orientationX/Z) and fixes the RNG block layout at 32x16
(src/main.cpp:115-116).  The same keys appear here as ``RunParams``; the
scene is rendered by ``patchmatch_cuda_amd.scene`` because the reference
ships no data.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass

BLUR_NONE, BLUR_GAUSS, BLUR_BOX = 0, 1, 2
INIT_DEPTH, INIT_INVERSE = 0, 1


@dataclass(frozen=True)
class SceneSpec:
    W: int
    H: int
    V: int
    n_planes: int
    n_spheres: int
    depth_min: float
    depth_max: float
    focal: float
    base_min: float
    base_max: float
    rot_max_deg: float
    noise_sigma: float
    scene_seed: int
    noise_seed: int


@dataclass(frozen=True)
class RunParams:
    half_window: int
    num_samples: int
    num_iterations: int
    depth_near: float
    depth_far: float
    spm_alpha: float = 0.45
    gaussian_sigma: float = 1.0
    blur_mode: int = BLUR_GAUSS
    init_mode: int = INIT_INVERSE
    orientation_x: float = 0.0
    orientation_z: float = 1.0
    block_x: int = 32
    block_y: int = 16
    rng_seed: int = 7

    def replace(self, **kw) -> "RunParams":
        return dataclasses.replace(self, **kw)


# 1 reference + 4 sources, 320x240, 6 planes, depth [2,10], f=300,
# baselines 0.2-0.5, yaw within 10 deg, window 11, 3 iterations.
C1_SMALL = (
    SceneSpec(320, 240, 4, 6, 0, 2.0, 10.0, 300.0, 0.2, 0.5, 10.0, 0.01, 1, 101),
    RunParams(half_window=5, num_samples=2, num_iterations=3, depth_near=2.0, depth_far=10.0),
)

# 1 reference + 10 sources, 2048x1536, 20 planes + 5 spheres, depth [1,20],
# f=1800, baselines 0.1-1.0, rotations within 15 deg, window 15, 4 iterations.
C2_REALISTIC = (
    SceneSpec(2048, 1536, 10, 20, 5, 1.0, 20.0, 1800.0, 0.1, 1.0, 15.0, 0.01, 2, 102),
    RunParams(half_window=7, num_samples=5, num_iterations=4, depth_near=1.0, depth_far=20.0),
)

# Precision + batching: 16 reference views, each with 10 sources, resolutions
# drawn uniformly from four sizes, FP16 texel storage, one batched run.
BATCH_SIZES = [(1024, 768), (1600, 1200), (2048, 1536), (3072, 2048)]


def c3_batch(n_views=16, seed=3):
    """The 16 views of the batched workload: (SceneSpec, RunParams) pairs.
    Each view is its own seeded world with the statistics of the 2048x1536
    workload (20 planes + 5 spheres, depths 1-20, baselines 0.1-1.0,
    rotations within 15 deg), the focal length scaled with the width (1800
    at 2048)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    _, rp = C2_REALISTIC
    out = []
    for i in range(n_views):
        W, H = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        out.append((SceneSpec(W, H, 10, 20, 5, 1.0, 20.0, 1800.0 * W / 2048.0, 0.1, 1.0, 15.0,
                              0.01, 300 + i, 400 + i), rp))
    return out


WORKLOADS = {
    "c1_small": C1_SMALL,
    "c2_realistic": C2_REALISTIC,
}


def tiny(W=48, H=36, V=2, hw=2, S=2, iters=2, seed=3, planes=3):
    """A small scene for quick parity checks."""
    return (
        SceneSpec(W, H, V, planes, 0, 2.0, 8.0, 0.9 * W, 0.2, 0.4, 8.0, 0.01, seed, seed + 100),
        RunParams(half_window=hw, num_samples=S, num_iterations=iters, depth_near=2.0, depth_far=8.0),
    )
