"""The reference-compatible C++ layer end to end on the GPU: pmx_cli (the
reference's main, src/main.cpp:10-128) reads an INI file and a Middlebury
camera list with 8-bit images, runs PatchMatch through the drop-in classes
and writes the depth map with saveToFile.  Its outputs must equal the C-ABI
engine's on the same inputs bit for bit, and agree with the oracle."""
import os
import subprocess
import sys
import types

import numpy as np
import pytest

from patchmatch_cuda_amd import capi, configs, formats as F, scene
from pmx_testlib import oracle_for, within

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _write_case(d, spec, rp):
    sc = scene.render(spec)
    imgs = [sc.ref] + list(sc.srcs)
    u8 = [np.clip(np.rint(im * 255.0), 0, 255).astype(np.uint8) for im in imgs]
    names = []
    for i, im in enumerate(u8):
        names.append(f"view{i:02d}.pgm")
        F.write_pgm(os.path.join(d, names[-1]), im)
    K, R, T = F.cams_to_krt(sc.cams)
    par = os.path.join(d, "scene_par.txt")
    F.write_middlebury(par, names, K, R, T)
    ini = os.path.join(d, "run.ini")
    F.write_ini(ini, {
        "imageInfoFileName": par, "refImageId": 0, "gpuId": 0,
        "halfWindowSize": rp.half_window, "numOfSamples": rp.num_samples,
        "SPMAlpha": rp.spm_alpha, "gaussianSigma": rp.gaussian_sigma,
        "numOfIterations": rp.num_iterations, "depthRangeNear": rp.depth_near,
        "depthRangeFar": rp.depth_far, "orientationX": 0, "orientationZ": 1,
        "outputFileName": os.path.join(d, "depth.bin"),
        "outputCostFileName": os.path.join(d, "cost.bin"),
    })
    return sc, u8, par, ini


def test_cli_equals_engine_and_oracle(tmp_path, gpu_ctx):
    d = str(tmp_path)
    spec, rp = configs.tiny(W=96, H=72, V=3, hw=3, S=2, iters=2)
    rp = rp.replace(init_mode=configs.INIT_DEPTH)   # the CLI's default is the reference's
    sc, u8, par, ini = _write_case(d, spec, rp)
    _say("case: pmx_cli on a 96x72 Middlebury list, 3 sources")
    r = subprocess.run([F.CLI, ini], capture_output=True, text=True, timeout=120)
    _say(r.stdout[-400:] + r.stderr[-400:])
    assert r.returncode == 0, r.stdout + r.stderr
    depth_cli = F.read_depth_file(os.path.join(d, "depth.bin"))
    cost_cli = F.read_depth_file(os.path.join(d, "cost.bin"))
    assert depth_cli.shape == (1, 72, 96) and cost_cli.shape == (3, 72, 96)

    # the same inputs through the C ABI: cameras exactly as the C++ reader holds them
    kind, c = F.c_read_cameras(par)
    assert kind == 2
    cams = np.concatenate([c["K"], c["R"], c["C"]], axis=1).astype(np.float64)
    eng = capi.Engine(gpu_ctx, rp, u8[0], u8[1:], cams, image_format=capi.IMG_U8)
    eng.prepare()
    d0, st0 = eng.depth(), eng.rng()
    eng.run()
    dg, cg = eng.depth(), eng.costs()
    assert np.array_equal(depth_cli[0], dg)
    assert np.array_equal(cost_cli, cg)

    pb = oracle_for(eng, rp, types.SimpleNamespace(cams=cams))
    do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples, rp.num_iterations)
    agree = within(dg, do, 0.01)
    _say(f"  within 1% of oracle: {agree:.4f}")
    eng.close()
    assert agree >= 0.99


def test_cli_rejects_window_like_the_reference(tmp_path):
    """hw outside 1..22: the reference's filterImage warning and exit(0)."""
    d = str(tmp_path)
    spec, rp = configs.tiny(W=48, H=36, V=2)
    _, _, _, ini = _write_case(d, spec, rp.replace(half_window=23))
    r = subprocess.run([F.CLI, ini], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "too large size of windows" in r.stdout
    assert not os.path.exists(os.path.join(d, "depth.bin"))
