"""pmx_blur_noalloc (no device allocation; the row pass in library-owned
scratch) against pmx_blur, through the C ABI: bit for bit, on two calls in a
row with the second smaller than the first (the scratch keeps the first
call's contents), in both filter modes; sizes beyond the scratch are
rejected before any launch."""
import ctypes
import sys

import numpy as np
import pytest

from patchmatch_cuda_amd import capi

pytestmark = pytest.mark.gpu

BLUR_GAUSS, BLUR_BOX = 1, 2


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _blur(ctx, fn, planes, sigma, mode):
    D, H, W = planes.shape
    a = capi.DeviceArray(ctx, (D * H, W), np.float32).upload(planes.reshape(D * H, W))
    try:
        ctx.check(getattr(capi.lib(), fn)(ctx.h, ctypes.c_void_p(a.ptr), W * 4, W, H, D,
                                          ctypes.c_float(sigma), mode))
        ctx.sync()
        return a.download().reshape(D, H, W)
    finally:
        a.close()


@pytest.mark.parametrize("mode,sigma", [(BLUR_GAUSS, 1.0), (BLUR_BOX, 1.7)])
def test_blur_noalloc_equals_blur(gpu_ctx, mode, sigma):
    rng = np.random.default_rng(17)
    for (D, H, W) in [(3, 1536, 2048), (2, 211, 333)]:      # the second call is smaller
        _say(f"case: pmx_blur_noalloc vs pmx_blur, mode {mode}, {D} planes of {W}x{H}")
        planes = rng.random((D, H, W), dtype=np.float32)
        want = _blur(gpu_ctx, "pmx_blur", planes, sigma, mode)
        got = _blur(gpu_ctx, "pmx_blur_noalloc", planes, sigma, mode)
        assert np.array_equal(got, want)
        assert not np.array_equal(got, planes)


def test_blur_noalloc_rejects_oversized_planes(gpu_ctx):
    _say("case: pmx_blur_noalloc beyond its scratch")
    a = capi.DeviceArray(gpu_ctx, (4, 4), np.float32)
    try:
        # 4097 x 2048 floats exceed the 2^23-float scratch: rejected on the host
        rc = capi.lib().pmx_blur_noalloc(gpu_ctx.h, ctypes.c_void_p(a.ptr), 4097 * 4, 4097, 2048,
                                         1, ctypes.c_float(1.0), BLUR_GAUSS)
        assert rc == 1                      # PMX_ERR_ARG
        assert "exceeds its scratch" in capi.last_error(gpu_ctx.h)
    finally:
        a.close()
