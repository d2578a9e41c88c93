"""GPU parity of FP16 texel storage (PMX_STORE_F16).

The images the NCC reads are rounded to FP16 (ties to even); homographies
stay FP64 and the window sums FP32.  Two questions are answered:

* parity: against the FP64 oracle fed the same rounded images, the FP16
  engine meets the FP32 engine's bars (costs within 1e-5, sweeps decision
  for decision, >= 99% of final depths within 1%);
* precision: against the FP64 oracle on the unrounded FP32 images, the
  error that FP16 storage itself adds is stated (printed) and bounded.
"""
import sys

import numpy as np
import pytest

from oracle import pmx_oracle as O
from patchmatch_cuda_amd import capi, configs

from pmx_testlib import (TINY, compare_sweep, f32, make_engine, oracle_for, rng_column_states,
                         to_frame, within)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _f16_engine(ctx, spec, rp):
    sc, eng = make_engine(ctx, spec, rp, prepare=False)
    eng.set_storage(capi.STORE_F16)
    eng.prepare()
    return sc, eng


@pytest.fixture(scope="module")
def tiny16(gpu_ctx):
    spec, rp = TINY
    sc, eng = _f16_engine(gpu_ctx, spec, rp)
    yield spec, rp, sc, eng
    eng.close()


@pytest.fixture(scope="module")
def c1_16(gpu_ctx):
    spec, rp = configs.C1_SMALL
    sc, eng = _f16_engine(gpu_ctx, spec, rp)
    yield spec, rp, sc, eng
    eng.close()


def test_storage_switch_and_rounding(gpu_ctx):
    """The rounded images equal numpy's float16 rounding of the FP32
    engine's pre-blurred images, bit for bit; switching back restores them."""
    spec, rp = TINY
    _say("case: fp16 rounding of the pre-blurred images (tiny)")
    sc, eng = make_engine(gpu_ctx, spec, rp)
    ref32, src32 = eng.images()
    assert eng.storage() == capi.STORE_F32
    eng.set_storage(capi.STORE_F16)
    with pytest.raises(capi.PmxError):
        eng.run()                       # a storage change needs a new prepare
    eng.prepare()
    ref16, src16 = eng.images()
    np.testing.assert_array_equal(ref16, ref32.astype(np.float16).astype(np.float32))
    np.testing.assert_array_equal(src16, src32.astype(np.float16).astype(np.float32))
    eng.set_storage(capi.STORE_F32)
    eng.prepare()
    r, s = eng.images()
    np.testing.assert_array_equal(r, ref32)
    np.testing.assert_array_equal(s, src32)
    with pytest.raises(capi.PmxError):
        eng.set_storage(7)
    eng.close()


def test_storage_round_trip_runs(gpu_ctx):
    """FP32 storage (the texel-pair block, allocated at the first FP32
    prepare), then FP16 (which frees it), then FP32 again (a new pair block):
    the two FP32 runs give the same depths, costs and generator states, bit
    for bit, and the FP16 run differs from them."""
    spec, rp = configs.C1_SMALL
    _say("case: storage FP32 -> FP16 -> FP32, full runs (C1)")
    sc, eng = make_engine(gpu_ctx, spec, rp, prepare=False)
    state = lambda: (eng.depth(), eng.costs(), eng.rng().view(np.uint32))  # noqa: E731
    try:
        eng.prepare()
        eng.run()
        a = state()
        eng.set_storage(capi.STORE_F16)
        eng.prepare()
        eng.run()
        h = state()
        eng.set_storage(capi.STORE_F32)
        eng.prepare()
        eng.run()
        b = state()
        for name, x, y in zip(("depth", "costs", "rng"), a, b):
            assert np.array_equal(x, y), f"second FP32 run: {name} differs"
        assert not np.array_equal(a[0], h[0])
    finally:
        eng.close()


def test_fp16_cost_all_views_within_1e5(c1_16):
    spec, rp, sc, eng = c1_16
    _say("case: fp16 cost all views c1")
    eng.prepare()
    depth = eng.depth()
    eng.cost_all_views(rotated=True)
    cg = eng.costs()
    pb = oracle_for(eng, rp, sc)
    co = np.swapaxes(pb.cost_all_views(to_frame(depth.astype(np.float64), True), rotated=True),
                     -1, -2)
    legit = np.abs(co) <= 2.0
    diff = np.where(legit, np.abs(cg - co), np.abs(cg - co) / np.maximum(np.abs(co), 1.0))
    _say(f"  vs oracle on the rounded images: max {diff.max():.3g} "
         f"p99.9 {np.quantile(diff, 0.999):.3g}")
    assert np.mean((cg == 2.0) == (co == 2.0)) == 1.0
    assert np.quantile(diff, 0.999) < 1e-5
    assert diff.max() < 1e-5
    # precision: the oracle on the FP32 (unrounded) images
    ref32 = O.blur(sc.ref, rp.gaussian_sigma, rp.blur_mode)
    src32 = np.stack([O.blur(s, rp.gaussian_sigma, rp.blur_mode) for s in sc.srcs])
    pb32 = O.Problem(ref32, src32, eng.homographies(), rp.half_window, f32(rp.depth_near),
                     f32(rp.depth_far), f32(rp.spm_alpha), O.inv_k4(sc.cams[0, :9]),
                     f32(rp.orientation_x), f32(rp.orientation_z))
    c32 = np.swapaxes(pb32.cost_all_views(to_frame(depth.astype(np.float64), True), rotated=True),
                      -1, -2)
    l32 = np.abs(c32) <= 2.0
    e = np.abs(cg - c32)[l32 & legit]
    _say(f"  FP16 storage error vs FP64 oracle on FP32 images: median {np.median(e):.3g} "
         f"p99 {np.quantile(e, 0.99):.3g} max {e.max():.3g}")
    assert np.quantile(e, 0.99) < 2e-3


@pytest.mark.parametrize("down,rotated", [(1, 1), (0, 0)])
def test_fp16_sweep_teacher_forced(tiny16, down, rotated):
    spec, rp, sc, eng = tiny16
    _say(f"case: fp16 sweep tiny down={down} rotated={rotated}")
    eng.prepare()
    eng.cost_all_views(rotated=True)
    eng.sweep(1, 1, 1)
    depth, costs, states = eng.depth(), eng.costs(), eng.rng()
    S = rp.num_samples
    pb = oracle_for(eng, rp, sc)
    d_s = to_frame(depth.astype(np.float64), rotated)
    c_s = to_frame(costs.astype(np.float64), rotated)
    st_o = states.copy()
    ids_o, tr = pb.sweep(d_s, c_s, st_o, down, rotated, S, record_ids=True, trace=True)
    eng.sweep(down, rotated, S, record_ids=True)
    res = compare_sweep(to_frame(eng.depth(), rotated), d_s, to_frame(eng.used_ids(), rotated),
                        ids_o, tr, down)
    _say(f"  compared {res['compared']}/{res['total']} mismatches {res['mismatches']}")
    assert res["mismatches"] == 0
    assert res["compared"] >= 0.6 * res["total"]
    clean = res["clean_cols"]
    Ws = d_s.shape[1]
    a = rng_column_states(eng.rng(), rotated, Ws)[clean].view(np.uint32)
    b = rng_column_states(st_o, rotated, Ws)[clean].view(np.uint32)
    assert clean.sum() > 0 and np.array_equal(a, b)


def test_fp16_full_run_c1(c1_16):
    """>= 99% of final depths within 1% of the oracle on the same (rounded)
    images; the agreement with the oracle on the FP32 images and with the
    ground truth is stated."""
    spec, rp, sc, eng = c1_16
    _say("case: fp16 full run c1")
    eng.prepare()
    d0, st0 = eng.depth(), eng.rng()
    eng.run()
    dg = eng.depth()
    pb = oracle_for(eng, rp, sc)
    do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples, rp.num_iterations)
    agree = within(dg, do, 0.01)
    ref32 = O.blur(sc.ref, rp.gaussian_sigma, rp.blur_mode)
    src32 = np.stack([O.blur(s, rp.gaussian_sigma, rp.blur_mode) for s in sc.srcs])
    pb32 = O.Problem(ref32, src32, eng.homographies(), rp.half_window, f32(rp.depth_near),
                     f32(rp.depth_far), f32(rp.spm_alpha), O.inv_k4(sc.cams[0, :9]),
                     f32(rp.orientation_x), f32(rp.orientation_z))
    d32, _, _ = pb32.run(d0.astype(np.float64), st0.copy(), rp.num_samples, rp.num_iterations)
    _say(f"  within 1%: of the oracle on rounded images {agree:.4f}; of the oracle on FP32 "
         f"images {within(dg, d32, 0.01):.4f}; ground truth fp16 {within(dg, sc.gt_depth):.4f} "
         f"fp32-oracle {within(d32, sc.gt_depth):.4f}")
    assert agree >= 0.99
    assert within(dg, sc.gt_depth) >= within(d32, sc.gt_depth) - 0.01
