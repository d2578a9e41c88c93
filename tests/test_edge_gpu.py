"""GPU parity at the edges of the input space, against the FP64 oracle fed
the engine's own images (teacher forcing):

* window sizes of every kernel shape: 15 (the static 15-row unroll), 21
  (32-lane hypothesis groups) and 45 (64-lane groups, the largest window
  filterImage accepts, src/array2d_refImg.cu:72-97), the last one taller
  than the image;
* sources of different sizes than the reference and each other (packed and
  zero-padded like copyData, src/patchMatch.cu:51-130);
* a single source view (the minimum, src/patchMatch.cu:138-142) and an
  image whose sides are not multiples of any block size;
* more than 16 source views or more than 7 draws per row (the general
  sweep kernels: the view CDF over the whole wave, draws resolved by one
  compare of floor(64 / V) draws at a time) and, above 32 views, the cost
  record transposes in 8 x 8 record tiles (src/patchMatch.cu:888-961 for the
  per-view messages and draws, which take any view count).

Bars as for the main parity tests: costs within 1e-5, a teacher-forced
sweep decision for decision, >= 99% of final depths within 1%."""
import sys

import numpy as np
import pytest

from patchmatch_cuda_amd import capi, configs, scene

from pmx_testlib import compare_sweep, oracle_for, to_frame, within

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(ctx, sc, eng, rp, tag):
    eng.prepare()
    depth = eng.depth()
    eng.cost_all_views(rotated=True)
    cg = eng.costs()
    pb = oracle_for(eng, rp, sc)
    co = np.swapaxes(pb.cost_all_views(to_frame(depth.astype(np.float64), True), rotated=True),
                     -1, -2)
    legit = np.abs(co) <= 2.0
    diff = np.where(legit, np.abs(cg - co), np.abs(cg - co) / np.maximum(np.abs(co), 1.0))
    assert np.mean((cg == 2.0) == (co == 2.0)) == 1.0
    # teacher-forced plain top-to-bottom sweep after one rotated sweep
    eng.sweep(1, 1, 1)
    d, c, st = eng.depth(), eng.costs(), eng.rng()
    pb = oracle_for(eng, rp, sc)
    d_s = d.astype(np.float64)
    ids_o, tr = pb.sweep(d_s, c.astype(np.float64), st.copy(), 1, 0, rp.num_samples,
                         record_ids=True, trace=True)
    eng.sweep(1, 0, rp.num_samples, record_ids=True)
    res = compare_sweep(eng.depth(), d_s, eng.used_ids(), ids_o, tr, 1)
    # full run
    eng.prepare()
    d0, st0 = eng.depth(), eng.rng()
    eng.run()
    do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples, rp.num_iterations)
    agree = within(eng.depth(), do, 0.01)
    _say(f"  {tag}: cost max {diff.max():.3g}; sweep compared {res['compared']}/{res['total']} "
         f"mismatches {res['mismatches']}; full run {agree:.4f} within 1%")
    assert diff.max() < 1e-5
    assert res["mismatches"] == 0
    assert res["compared"] >= 0.5 * res["total"]
    assert agree >= 0.99


@pytest.mark.parametrize("hw", [7, 10, 22])
def test_window_sizes(gpu_ctx, hw):
    _say(f"case: window {2 * hw + 1}x{2 * hw + 1} on 64x48, 2 sources")
    spec, rp = configs.tiny(W=64, H=48, V=2, hw=hw, S=2, iters=2, seed=31 + hw)
    sc = scene.render(spec)
    eng = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs), sc.cams)
    _check(gpu_ctx, sc, eng, rp, f"hw={hw}")
    eng.close()


def test_ragged_sources(gpu_ctx):
    """Sources cropped from the top-left (the cameras stay valid) to sizes
    below the reference's and each other's."""
    _say("case: sources of 64x48, 54x42 and 61x39 against a 64x48 reference")
    spec, rp = configs.tiny(W=64, H=48, V=3, hw=3, S=2, iters=2, seed=41)
    sc = scene.render(spec)
    srcs = [sc.srcs[0], np.ascontiguousarray(sc.srcs[1][:42, :54]),
            np.ascontiguousarray(sc.srcs[2][:39, :61])]
    eng = capi.Engine(gpu_ctx, rp, sc.ref, srcs, sc.cams)
    assert (eng.maxW, eng.maxH) == (64, 48)
    eng.prepare()
    ref, src_imgs = eng.images()
    # zero outside each source's own extent
    assert not src_imgs[1][42:, :].any() and not src_imgs[1][:, 54:].any()
    assert not src_imgs[2][39:, :].any() and not src_imgs[2][:, 61:].any()
    _check(gpu_ctx, sc, eng, rp, "ragged")
    eng.close()


def test_single_source_and_odd_size(gpu_ctx):
    _say("case: one source view, 33x17 image")
    spec, rp = configs.tiny(W=33, H=17, V=1, hw=2, S=1, iters=2, seed=51)
    sc = scene.render(spec)
    eng = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs), sc.cams)
    _check(gpu_ctx, sc, eng, rp, "V=1 33x17")
    eng.close()


@pytest.mark.parametrize("V,S,hw", [(5, 12, 2), (20, 9, 3), (40, 3, 7)])
def test_many_views_and_draws(gpu_ctx, V, S, hw):
    _say(f"case: {V} sources, S = {S}, window {2 * hw + 1}x{2 * hw + 1} on 48x40")
    spec, rp = configs.tiny(W=48, H=40, V=V, hw=hw, S=S, iters=2, seed=61 + V)
    sc = scene.render(spec)
    eng = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs), sc.cams)
    _check(gpu_ctx, sc, eng, rp, f"V={V} S={S}")
    eng.close()
