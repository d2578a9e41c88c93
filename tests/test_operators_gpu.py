"""GPU parity of the stand-alone operators of the C ABI against the oracle,
called through the C ABI with device buffers:

  pmx_transpose_f32         cudaTranspose::transpose2dData (src/cudaTranspose.h:52-71)
  pmx_blur                  GaussianBlurCUDA::FilterMultipleImages (src/GaussianBlurCUDA.cu:189-199)
  pmx_box_sums              Array2d_refImg::filterImage (src/array2d_refImg.cu:70-131)
  pmx_rng_init / _fill_depth  Array2D_psng + randNumGen (src/array2d_psng.cu:3-31,
                            src/array2d_wrapper.cu:3-20)
  pmx_compose_homographies  Image::init_relative (src/image.cpp:4-13)

Transposes and generator states are compared bit for bit."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import pmx_oracle as O
from patchmatch_cuda_amd import capi, configs, scene

pytestmark = pytest.mark.gpu
_vp = ctypes.c_void_p


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("W,H", [(1, 1), (31, 33), (33, 31), (257, 129)])
def test_transpose_is_exact(gpu_ctx, W, H):
    _say(f"case: transpose {W}x{H}")
    rng = np.random.default_rng(W * 1000 + H)
    a = rng.standard_normal((H, W)).astype(np.float32)
    src = capi.DeviceArray(gpu_ctx, (H, W), np.float32).upload(a)
    dst = capi.DeviceArray(gpu_ctx, (W, H), np.float32)
    gpu_ctx.check(capi.lib().pmx_transpose_f32(gpu_ctx.h, _vp(src.ptr), _vp(dst.ptr), W, H,
                                               W * 4, H * 4))
    out = dst.download()
    assert np.array_equal(out.view(np.uint32), O.transpose(a).astype(np.float32).view(np.uint32))
    src.close()
    dst.close()


@pytest.mark.parametrize("mode,sigma", [(configs.BLUR_GAUSS, 1.0), (configs.BLUR_GAUSS, 2.3),
                                        (configs.BLUR_BOX, 1.0), (configs.BLUR_NONE, 1.0)])
def test_blur_of_stacked_planes(gpu_ctx, mode, sigma):
    _say(f"case: blur mode {mode} sigma {sigma}, 3 stacked 61x37 planes")
    W, H, D = 61, 37, 3
    rng = np.random.default_rng(7)
    planes = rng.uniform(0, 1, (D, H, W)).astype(np.float32)
    buf = capi.DeviceArray(gpu_ctx, (D * H, W), np.float32).upload(planes.reshape(D * H, W))
    gpu_ctx.check(capi.lib().pmx_blur(gpu_ctx.h, _vp(buf.ptr), W * 4, W, H, D, sigma, mode))
    out = buf.download().reshape(D, H, W)
    for k in range(D):
        np.testing.assert_allclose(out[k], O.blur(planes[k], sigma, mode), atol=2e-6)
    if mode == configs.BLUR_NONE:
        assert np.array_equal(out, planes)
    buf.close()


@pytest.mark.parametrize("hw", [1, 5, 22])
def test_box_sums(gpu_ctx, hw):
    _say(f"case: window sums hw={hw} on 50x41")
    W, H = 50, 41
    img = np.random.default_rng(hw).uniform(0, 1, (H, W)).astype(np.float32)
    src = capi.DeviceArray(gpu_ctx, (H, W), np.float32).upload(img)
    oI, oS, oSS = (capi.DeviceArray(gpu_ctx, (H, W), np.float32) for _ in range(3))
    gpu_ctx.check(capi.lib().pmx_box_sums(gpu_ctx.h, _vp(src.ptr), W * 4, W, H, hw, _vp(oI.ptr),
                                          _vp(oS.ptr), _vp(oSS.ptr), W * 4))
    sI, sII = O.box_sums(img.astype(np.float64), hw)
    assert np.array_equal(oI.download(), img)
    np.testing.assert_allclose(oS.download(), sI, rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(oSS.download(), sII, rtol=1e-6, atol=1e-5)
    # the accepted range (src/array2d_refImg.cu:72-97), checked before any launch
    with pytest.raises(capi.PmxError):
        gpu_ctx.check(capi.lib().pmx_box_sums(gpu_ctx.h, _vp(src.ptr), W * 4, W, H, 23,
                                              _vp(oI.ptr), _vp(oS.ptr), _vp(oSS.ptr), W * 4))
    for a in (src, oI, oS, oSS):
        a.close()


@pytest.mark.parametrize("mode", [configs.INIT_DEPTH, configs.INIT_INVERSE])
def test_generator_and_initial_depth(gpu_ctx, mode):
    _say(f"case: generator states and initial depth, mode {mode}, 70x45")
    W, H, near, far = 70, 45, 2.0, 8.0
    st = capi.DeviceArray(gpu_ctx, (H, W, 6), np.uint32)
    d = capi.DeviceArray(gpu_ctx, (H, W), np.float32)
    gpu_ctx.check(capi.lib().pmx_rng_init(gpu_ctx.h, _vp(st.ptr), W * 24, W, H, 32, 16, 11))
    states0 = st.download().copy()
    gpu_ctx.check(capi.lib().pmx_rng_fill_depth(gpu_ctx.h, _vp(st.ptr), W * 24, _vp(d.ptr), W * 4,
                                                W, H, near, far, mode))
    so = O.rng_init_grid(W, H, 32, 16, 11)
    assert np.array_equal(states0.reshape(H, W * 6), so.view(np.uint32).reshape(H, W * 6))
    do = O.init_depth(so, near, far, mode)   # advances so as the fill does
    np.testing.assert_allclose(d.download(), do, rtol=2e-6)
    assert np.array_equal(st.download().reshape(H, W * 6), so.view(np.uint32).reshape(H, W * 6))
    st.close()
    d.close()


def test_compose_homographies(gpu_ctx):
    _say("case: homography composition, 5 views")
    spec, rp = configs.tiny(W=64, H=48, V=5, seed=13)
    sc = scene.render(spec)
    ref = capi.Camera.from_row(sc.cams[0])
    srcs = (capi.Camera * 5)(*[capi.Camera.from_row(sc.cams[v + 1]) for v in range(5)])
    out = np.zeros(5 * 18)
    gpu_ctx.check(capi.lib().pmx_compose_homographies(gpu_ctx.h, ctypes.byref(ref), srcs, 5,
                                                      0.25, 0.75, _vp(out.ctypes.data)))
    want = O.homographies(sc.cams, 0.25, 0.75)
    np.testing.assert_allclose(out.reshape(want.shape), want, rtol=1e-12,
                               atol=1e-12 * np.abs(want).max())


def test_engine_without_pre_blur_reads_the_input(gpu_ctx):
    """blurMode none (the reference's own behaviour: its filter is never
    called): the images the NCC reads are the uploaded ones."""
    _say("case: engine with blur mode none")
    spec, rp = configs.tiny(W=40, H=30, V=2, seed=17)
    rp = rp.replace(blur_mode=configs.BLUR_NONE)
    sc = scene.render(spec)
    eng = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs), sc.cams)
    eng.prepare()
    ref, srcs = eng.images()
    assert np.array_equal(ref, sc.ref)
    assert np.array_equal(srcs, sc.srcs)
    eng.close()
