"""The reference's helper classes through the drop-in C++ layer
(libpmx_drop_in.so), driven by the C++ program tests/dropin_helpers.cpp and
checked against the FP64 oracle:

* Array2D_psng (src/array2d_psng.cu:3-31): generator states bit for bit;
* Array2D_wrapper<float>::randNumGen (src/array2d_wrapper.cu:3-20): the
  initial depth draw and the states after it;
* Array2d_refImg::filterImage (src/array2d_refImg.cu:70-131): I = u/255 and
  the window sums of I and I^2 with zero border;
* cudaTranspose::transpose2dData (src/cudaTranspose.h:52-71) on odd sizes;
* GaussianBlurCUDA::FilterMultipleImages (src/GaussianBlurCUDA.cu:189-199)
  with Gaussian and the reference's uniform taps;
* CudaArray_wrapper::array3DCopy (src/cudaArray_wrapper.h:27-60) round trips;
* PatchMatch::_refImage and _psngState (src/patchMatch.h:41, :49) after the
  constructor: the pre-blurred reference, its window sums, and the states
  after the initial-depth draw; after runPatchMatch only the row-0 and
  column-0 states the sweeps use have moved."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import pmx_oracle as O
from patchmatch_cuda_amd import configs, formats as F, scene

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PROG = os.path.join(HERE, "dropin_helpers")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_helper_classes_against_oracle(tmp_path):
    assert os.path.exists(PROG), "tests/dropin_helpers is built by __graft_entry__.build()"
    W, H, hw, bx, by, seed = 45, 37, 3, 32, 16, 7
    near, far, sigma = np.float32(2.0), np.float32(8.0), np.float32(1.0)
    rng = np.random.default_rng(5)
    img8 = rng.integers(0, 256, (H, W), dtype=np.uint8)
    planes = rng.random((2, H, W), dtype=np.float32)
    Wt, Ht = 33, 31
    tin = rng.random((Ht, Wt), dtype=np.float32)
    spec, rp = configs.tiny(W=W, H=H, V=2, hw=hw, S=2, iters=2, seed=9)
    sc = scene.render(spec)
    K, R, T = F.cams_to_krt(sc.cams)
    imgs = [sc.ref] + list(sc.srcs)
    inp = os.path.join(str(tmp_path), "in.bin")
    out = os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as f:
        f.write(np.array([W, H, hw, bx, by], np.int32).tobytes())
        f.write(np.array([seed], np.uint32).tobytes())
        f.write(np.array([near, far, sigma], np.float32).tobytes())
        f.write(img8.tobytes())
        f.write(planes.tobytes())
        f.write(np.array([Wt, Ht], np.int32).tobytes())
        f.write(tin.tobytes())
        f.write(np.array([len(imgs), 2, 2], np.int32).tobytes())
        f.write(np.array([0.45], np.float32).tobytes())
        for i, im in enumerate(imgs):
            for m in (K[i], R[i], T[i]):
                f.write(np.asarray(m, np.float32).ravel().tobytes())
            f.write(np.ascontiguousarray(im, np.float32).tobytes())
    _say("case: drop-in helper classes through libpmx_drop_in.so")
    r = subprocess.run([PROG, inp, out], capture_output=True, text=True, timeout=120)
    _say(r.stdout[-300:] + r.stderr[-300:])
    assert r.returncode == 0
    buf = open(out, "rb").read()
    pos = 0

    def take(dtype, shape):
        nonlocal pos
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        a = np.frombuffer(buf[pos:pos + n], dtype).reshape(shape)
        pos += n
        return a

    st_dt = O.RNG_DTYPE
    st0 = take(st_dt, (H, W))
    depth = take(np.float32, (H, W))
    st1 = take(st_dt, (H, W))
    I, sI, sII = (take(np.float32, (H, W)) for _ in range(3))
    tout = take(np.float32, (Wt, Ht))
    blur_g = take(np.float32, (2, H, W))
    blur_b = take(np.float32, (2, H, W))
    rt32 = take(np.float32, (2, H, W))
    rt8 = take(np.uint8, (H, W))
    pI, psI, psII = (take(np.float32, (H, W)) for _ in range(3))
    pst = take(st_dt, (H, W))
    pdepth = take(np.float32, (H, W))
    pst_run = take(st_dt, (H, W))
    assert pos == len(buf)

    want0 = O.rng_init_grid(W, H, bx, by, seed)
    assert np.array_equal(st0.view(np.uint32), want0.view(np.uint32))
    st_o = want0.copy()
    d_o = O.init_depth(st_o, float(near), float(far), configs.INIT_DEPTH)
    np.testing.assert_allclose(depth, d_o, rtol=2e-6)
    assert np.array_equal(st1.view(np.uint32), st_o.view(np.uint32))

    want_I = img8.astype(np.float32) / np.float32(255.0)
    assert np.array_equal(I, want_I)
    oI, oII = O.box_sums(want_I.astype(np.float64), hw)
    np.testing.assert_allclose(sI, oI, rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(sII, oII, rtol=1e-6, atol=1e-5)

    assert np.array_equal(tout, tin.T)

    for got, mode in ((blur_g, configs.BLUR_GAUSS), (blur_b, configs.BLUR_BOX)):
        for k in range(2):
            np.testing.assert_allclose(got[k], O.blur(planes[k], float(sigma), mode), atol=2e-6)
    assert not np.allclose(blur_g, blur_b)

    assert np.array_equal(rt32, planes) and np.array_equal(rt8, img8)

    # PatchMatch members: the constructor pre-blurs (Gaussian, sigma) and
    # draws the initial depth from the generator (seed 7, 32 x 16 blocks)
    np.testing.assert_allclose(pI, O.blur(sc.ref, float(sigma), configs.BLUR_GAUSS), atol=2e-6)
    oI, oII = O.box_sums(pI.astype(np.float64), hw)
    np.testing.assert_allclose(psI, oI, rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(psII, oII, rtol=1e-6, atol=1e-5)
    st_p = O.rng_init_grid(W, H, 32, 16, 7)
    O.init_depth(st_p, float(near), float(far), configs.INIT_DEPTH)
    assert np.array_equal(pst.view(np.uint32), st_p.view(np.uint32))
    moved = np.any(pst_run.view(np.uint32).reshape(H, W, 6) != pst.view(np.uint32).reshape(H, W, 6),
                   axis=-1)
    assert moved[0].all() and moved[:, 0].all() and not moved[1:, 1:].any()
    assert np.all((pdepth >= near) & (pdepth <= far))
    _say("  states, draws, window sums, transpose, blur, round trips and PatchMatch views agree")
