"""GPU parity at the realistic workloads, with the FP64 oracle as the checker.

* 2048x1536 plain (1 reference + 10 sources, window 15, S = 5, 4
  iterations): the centred 768x576 crop of the reference against all ten
  full-size sources.  That parameter set is the reference's run<15> path
  (src/patchMatch.cu:274-303, :323-394): ten views, five draws deduplicated
  over ten lanes, and more hypotheses per row than one pass of lane groups
  holds.  Costs within 1e-5, teacher-forced sweeps in both frames and both
  launch shapes decision for decision, >= 99% of a full run within 1%.
* the 16-view mixed-resolution batch with FP16 texels (one persistent
  launch per sweep step, replayed from a hipGraph): every view equals its
  standalone run bit for bit, and two views (one 3072x2048) are checked
  against the oracle on 512x384 crops.
* the learned proposal at 2048x1536 with the packaged weights: the learned
  run on a crop against the oracle running the same learned algorithm, and
  the trained network's measured effect.
* the serving loop on a short 4 requests/s stream with the learned proposal
  on: one served depth map against the oracle.
"""
import os
import sys
import time

import numpy as np
import pytest

from oracle import pmx_oracle as O
from patchmatch_cuda_amd import capi, configs, proposal, scene, serving

from pmx_testlib import (compare_sweep, centred_crop, f32, ncc_diff, oracle_for, oracle_problem,
                         proposal_planes, rng_column_states, to_frame, within)

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _threads():
    old = O.get_threads()
    O.set_threads(min(ORACLE_THREADS, os.cpu_count() or 1))
    yield
    O.set_threads(old)


@pytest.fixture(scope="module")
def c2_scene():
    spec, rp = configs.C2_REALISTIC
    t = time.perf_counter()
    sc = scene.render(spec)
    _say(f"rendered 2048x1536 with 10 sources in {time.perf_counter() - t:.1f} s")
    return spec, rp, sc


@pytest.fixture(scope="module")
def c2_crop(gpu_ctx, c2_scene):
    spec, rp, sc = c2_scene
    (x0, y0), (ref, srcs, cams, gt) = centred_crop(sc, 768, 576)
    eng = capi.Engine(gpu_ctx, rp, ref, list(srcs), cams)
    eng.prepare()
    yield rp, (ref, srcs, cams, gt), eng
    eng.close()


def _oracle_of(eng, rp, cams):
    ref, srcs = eng.images()
    return oracle_problem(ref, srcs, eng.homographies(), cams[0], rp)


def test_c2_costs_within_1e5(c2_crop):
    rp, (ref, srcs, cams, gt), eng = c2_crop
    _say("case: 2048x1536 workload, 768x576 crop, cost of all views (V=10, window 15)")
    eng.prepare()
    depth = eng.depth()
    eng.cost_all_views(rotated=True)
    cg = eng.costs()
    pb = _oracle_of(eng, rp, cams)
    co = np.swapaxes(pb.cost_all_views(to_frame(depth.astype(np.float64), True), rotated=True),
                     -1, -2)
    diff = ncc_diff(cg, co)
    _say(f"  max |dcost| {diff.max():.3g}  p99.9 {np.quantile(diff, 0.999):.3g}; "
         f"sentinels {int((co == 2.0).sum())}")
    assert np.mean((cg == 2.0) == (co == 2.0)) == 1.0
    assert diff.max() < 1e-5


@pytest.mark.parametrize("waves", [1, 2, 4])
@pytest.mark.parametrize("down,rotated", [(1, 0), (0, 1)])
def test_c2_sweep_teacher_forced(c2_crop, waves, down, rotated):
    """S = 5 on ten views: up to ten phase-B hypotheses per row, spread over
    several passes of four lane groups (one wave per column) or combined
    across two or four waves."""
    rp, (ref, srcs, cams, gt), eng = c2_crop
    _say(f"case: 2048x1536 workload crop, sweep down={down} rotated={rotated} waves={waves}")
    eng.set_sweep_waves(waves)
    try:
        eng.prepare()
        eng.cost_all_views(rotated=True)
        eng.sweep(1, 1, 1)                      # iteration 0's first sweep (S = 1)
        eng.sweep(1, 0, rp.num_samples)         # a first S = 5 sweep
        depth, costs, states = eng.depth(), eng.costs(), eng.rng()
        pb = _oracle_of(eng, rp, cams)
        d_s = to_frame(depth.astype(np.float64), rotated)
        c_s = to_frame(costs.astype(np.float64), rotated)
        st_o = states.copy()
        ids_o, tr = pb.sweep(d_s, c_s, st_o, down, rotated, rp.num_samples, record_ids=True,
                             trace=True)
        eng.sweep(down, rotated, rp.num_samples, record_ids=True)
        dg = to_frame(eng.depth(), rotated)
        ids_g = to_frame(eng.used_ids(), rotated)
        res = compare_sweep(dg, d_s, ids_g, ids_o, tr, down)
        _say(f"  compared {res['compared']}/{res['total']} mismatches {res['mismatches']} "
             f"(depth {res['depth_mismatch']}, ids {res['ids_mismatch']}) first {res['first_bad']}")
        assert res["mismatches"] == 0
        # columns of 576-768 rows reach a near-tie decision sooner than short ones
        assert res["compared"] >= 0.25 * res["total"]
        clean = res["clean_cols"]
        Ws = d_s.shape[1]
        a = rng_column_states(eng.rng(), rotated, Ws)[clean].view(np.uint32)
        b = rng_column_states(st_o, rotated, Ws)[clean].view(np.uint32)
        assert clean.sum() > 0 and np.array_equal(a, b)
        # sweeps with more than one pass of hypotheses per row were exercised
        choice = tr["choice"]
        assert np.mean(choice[choice != 255] != 1) > 0.05
    finally:
        eng.set_sweep_waves(0)


def test_c2_full_run_within_1pct(c2_crop):
    rp, (ref, srcs, cams, gt), eng = c2_crop
    _say("case: 2048x1536 workload crop, full run (4 iterations, S = 5)")
    eng.prepare()
    d0, st0 = eng.depth(), eng.rng()
    eng.run()
    dg = eng.depth()
    pb = _oracle_of(eng, rp, cams)
    t = time.perf_counter()
    do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples, rp.num_iterations,
                      record_ids=False)
    agree = within(dg, do, 0.01)
    _say(f"  oracle {time.perf_counter() - t:.1f} s; within 1% of the oracle {agree:.4f}; "
         f"of ground truth gpu {within(dg, gt):.4f} oracle {within(do, gt):.4f}")
    assert agree >= 0.99


# ------------------------------------------------------------ 16-view batch
@pytest.fixture(scope="module")
def c3(gpu_ctx):
    from concurrent.futures import ThreadPoolExecutor

    views = configs.c3_batch()
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=4) as pool:
        scenes = list(pool.map(lambda v: scene.render(v[0]), views))
    _say(f"rendered the 16 views in {time.perf_counter() - t:.1f} s: "
         f"{[f'{v[0].W}x{v[0].H}' for v in views]}")
    yield views, scenes


def test_c3_batch_equals_standalone(gpu_ctx, c3):
    views, scenes = c3
    _say("case: 16-view mixed-resolution batch, FP16, graph, launch shape 3")
    engs = []
    for (spec, rp), sc in zip(views, scenes):
        e = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs), sc.cams)
        e.set_storage(capi.STORE_F16)
        engs.append(e)
    try:
        b = capi.Batch(gpu_ctx, engs)
        b.set_launch(3, 0)
        b.set_graph(True)
        got = []
        for rep in range(2):            # capture, then a replay
            b.prepare()
            ms = b.run(timed=True)
            got = [(e.depth(), e.used_ids(), e.rng().view(np.uint32)) for e in engs]
            _say(f"  batch run {rep}: {ms:.1f} ms")
        b.close()
        for i, e in enumerate(engs):
            e.set_sweep_waves(1)
            e.prepare()
            e.run()
            for name, x, y in zip(("depth", "ids", "rng"), got[i],
                                  (e.depth(), e.used_ids(), e.rng().view(np.uint32))):
                assert np.array_equal(x, y), f"view {i} ({e.W}x{e.H}): {name} differs"
            gt = scenes[i].gt_depth
            _say(f"  view {i} {e.W}x{e.H}: equal to its standalone run; "
                 f"{within(got[i][0], gt):.4f} within 1% of ground truth")
    finally:
        for e in engs:
            e.close()


def test_c3_views_against_oracle(gpu_ctx, c3):
    """Two views of the batch (a 3072x2048 one and the first of another
    size) on their centred 512x384 crops, batched with FP16 texels, against
    the oracle on the same FP16-rounded images (the parity bar) and on the
    FP32 images (the error FP16 storage adds, reported)."""
    views, scenes = c3
    big = next(i for i, v in enumerate(views) if (v[0].W, v[0].H) == (3072, 2048))
    other = next(i for i, v in enumerate(views) if (v[0].W, v[0].H) != (3072, 2048))
    crops, engs = [], []
    for i in (big, other):
        spec, rp = views[i]
        (x0, y0), (ref, srcs, cams, gt) = centred_crop(scenes[i], 512, 384)
        e = capi.Engine(gpu_ctx, rp, ref, list(srcs), cams)
        e.set_storage(capi.STORE_F16)
        crops.append((i, rp, ref, srcs, cams, gt))
        engs.append(e)
    try:
        b = capi.Batch(gpu_ctx, engs)
        b.set_launch(3, 0)
        b.set_graph(True)
        b.prepare()
        init = [(e.depth(), e.rng()) for e in engs]
        b.run()
        b.close()
        for (i, rp, ref, srcs, cams, gt), e, (d0, st0) in zip(crops, engs, init):
            dg = e.depth()
            r16, s16 = e.images()
            pb = oracle_problem(r16, s16, e.homographies(), cams[0], rp)
            do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples,
                              rp.num_iterations, record_ids=False)
            agree = within(dg, do)
            ref32 = O.blur(ref, rp.gaussian_sigma, rp.blur_mode)
            srcs32 = np.stack([O.blur(s, rp.gaussian_sigma, rp.blur_mode) for s in srcs])
            pb32 = oracle_problem(ref32, srcs32, e.homographies(), cams[0], rp)
            d32, _, _ = pb32.run(d0.astype(np.float64), st0.copy(), rp.num_samples,
                                 rp.num_iterations, record_ids=False)
            _say(f"  view {i} ({views[i][0].W}x{views[i][0].H}) 512x384 crop: {agree:.4f} within "
                 f"1% of the oracle on the FP16 images; {within(dg, d32):.4f} of the oracle on "
                 f"FP32 images; ground truth gpu {within(dg, gt):.4f} fp32 oracle "
                 f"{within(d32, gt):.4f}")
            assert agree >= 0.99
            assert abs(within(dg, gt) - within(d32, gt)) < 0.02
    finally:
        for e in engs:
            e.close()


# ----------------------------------------------------------- learned proposal
def test_c4_learned_against_oracle(gpu_ctx, c2_scene):
    """The packaged network on the 2048x1536 view: its proposal planes, the
    learned run on the centred 640x480 crop against the oracle running the
    same learned algorithm (fourth candidate on the same planes and hash
    draws), and what the trained network measurably provides."""
    spec, rp, sc = c2_scene
    _say("case: learned proposal, 2048x1536 with 10 sources, packaged weights")
    tau = 0.2
    eng = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs), sc.cams)
    W0, H0 = proposal.low_res(eng.W, eng.H)
    net = proposal.Net(gpu_ctx, 1, H0, W0, rp.depth_near, rp.depth_far)
    net.load(proposal.DEFAULT_WEIGHTS)
    try:
        eng.prepare()
        d_init = eng.depth()
        prop, spread = proposal_planes(gpu_ctx, eng, net, tau)
        gt = sc.gt_depth
        cover = prop > 0
        rel = np.abs(prop - gt) / gt
        net5 = float(np.mean(rel[cover] <= 0.05))
        init5 = float(np.mean(np.abs(d_init - gt) / gt <= 0.05))
        _say(f"  proposal on {cover.mean():.3f} of pixels; within 5% of ground truth: "
             f"{net5:.3f} of them (random initial depth: {init5:.3f})")
        # the trained network is far better than the initial depths it
        # would replace, on most of the image
        assert cover.mean() > 0.5
        assert net5 > 3 * init5 and net5 > 0.2
        # learned run on the crop against the oracle
        w, h = 640, 480
        (x0, y0), (ref, srcs, cams, gtc) = centred_crop(sc, w, h)
        pc = np.ascontiguousarray(prop[y0:y0 + h, x0:x0 + w])
        sc_ = np.ascontiguousarray(spread[y0:y0 + h, x0:x0 + w])
        ce = capi.Engine(gpu_ctx, rp, ref, list(srcs), cams)
        ce.set_proposal_seeding(False)
        pd = capi.DeviceArray(gpu_ctx, (h, w), np.float32).upload(pc)
        sd = capi.DeviceArray(gpu_ctx, (h, w), np.float32).upload(sc_)
        try:
            res = {}
            for learned in (False, True):
                ce.prepare()
                if learned:
                    ce.set_proposal(pd.ptr, w * 4, sd.ptr)
                d0, st0 = ce.depth(), ce.rng()
                ce.run()
                res[learned] = ce.depth()
                ce.set_proposal(None, 0)
            pb = _oracle_of(ce, rp, cams)
            pb.set_proposal(pc.astype(np.float64), sc_.astype(np.float64))
            do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples,
                              rp.num_iterations, record_ids=False)
            agree = within(res[True], do)
            _say(f"  640x480 crop: learned GPU {agree:.4f} within 1% of the learned oracle; "
                 f"ground truth learned {within(res[True], gtc):.4f} plain "
                 f"{within(res[False], gtc):.4f}")
            assert agree >= 0.99
            # the candidate never costs accuracy: NCC stays the judge
            assert within(res[True], gtc) >= within(res[False], gtc) - 0.005
        finally:
            pd.close()
            sd.close()
            ce.close()
    finally:
        net.close()
        eng.close()


# --------------------------------------------------------------- serving loop
def test_c5_serving_stream_learned(gpu_ctx):
    """A 4 requests/s Poisson stream (3 s of arrivals) of the serving mix
    (640x480 to 3072x2048, 4-20 sources, FP16 texels) with the learned
    proposal on.  Every request completes; the smallest served request is
    checked against the oracle running the same learned algorithm on the
    same FP16 images, generator states and proposal planes."""
    _say("case: serving loop, 4 requests/s for 3 s, learned proposal")
    _, rp = configs.C2_REALISTIC
    t = time.perf_counter()
    pool = serving.scene_pool()
    _say(f"  scene pool rendered in {time.perf_counter() - t:.1f} s")
    seed = 11
    while not any(r.size == 0 for r in serving.make_stream(4.0, 3.0, seed)):
        seed += 1
    stream = serving.make_stream(4.0, 3.0, seed)
    ctx2 = capi.Context(0)
    try:
        srv = serving.Server([gpu_ctx, ctx2], rp, pool, 64 * 2 ** 30, proposal.DEFAULT_WEIGHTS,
                             rp.num_iterations, 0.2)
        srv.keep_results = True
        res = srv.run(stream, log=_say)
        _say(f"  {res['requests']} requests, {res['throughput_req_s']:.2f} req/s, "
             f"p50 {res['latency_s']['p50']:.2f} s, p99 {res['latency_s']['p99']:.2f} s, "
             f"groups {res['groups_by_launch_shape']}")
        assert res["requests"] == len(stream)
        assert srv.resident == 0
        req = min((r for r in stream if r.size == 0), key=lambda r: r.n_src)
        sc = pool[req.size]
        served = srv.results[req.rid]
        # the same request standalone: images, generator states, proposal
        e = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs[:req.n_src]), sc.cams[:req.n_src + 1])
        e.set_storage(capi.STORE_F16)
        e.set_sweep_waves(1)
        e.set_proposal_seeding(False)
        e.set_proposal_sweeps(srv.proposal_sweeps)
        W0, H0 = proposal.low_res(e.W, e.H)
        net = proposal.Net(gpu_ctx, 1, H0, W0, rp.depth_near, rp.depth_far)
        net.load(proposal.DEFAULT_WEIGHTS)
        try:
            e.prepare()
            d0, st0 = e.depth(), e.rng()
            prop, spread = proposal_planes(gpu_ctx, e, net, 0.2)
            e.propose(net, tau=0.2)
            e.run()
            assert np.array_equal(e.depth(), served)        # served == standalone, bit for bit
            r16, s16 = e.images()
            pb = oracle_problem(r16, s16, e.homographies(), sc.cams[0], rp)
            pb.set_proposal(prop.astype(np.float64), spread.astype(np.float64))
            pb.set_proposal_sweeps(srv.proposal_sweeps)
            do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples,
                              rp.num_iterations, record_ids=False)
            agree = within(served, do)
            _say(f"  request {req.rid} ({e.W}x{e.H}, {req.n_src} sources): {agree:.4f} of served "
                 f"depths within 1% of the oracle")
            assert agree >= 0.99
        finally:
            net.close()
            e.close()
    finally:
        ctx2.close()
