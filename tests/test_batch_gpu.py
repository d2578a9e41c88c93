"""GPU tests of batched runs (pmx_batch_*): several reference views of
different sizes and view counts in one persistent sweep launch per sweep
step, replayed from a hipGraph.

Every view's sweep columns are independent of the other views', so each
engine of a batch must produce exactly what a run of that engine alone
produces with the same waves per column strip: depths, costs, view ids and
generator states bit for bit.  The standalone run is itself pinned to the
FP64 oracle (test_gpu_parity.py); one batch member is also checked against
the oracle directly."""
import sys

import numpy as np
import pytest

from patchmatch_cuda_amd import capi, configs

from pmx_testlib import make_engine, oracle_for, within

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# (W, H, V, seed): landscape, portrait and odd sizes, different view counts
SHAPES = [(48, 36, 2, 3), (30, 44, 3, 4), (57, 23, 2, 5), (40, 40, 4, 6)]


def _engines(ctx, storage=capi.STORE_F32, shapes=SHAPES):
    out = []
    for W, H, V, seed in shapes:
        spec, rp = configs.tiny(W=W, H=H, V=V, hw=2, S=2, iters=2, seed=seed)
        sc, eng = make_engine(ctx, spec, rp, prepare=False)
        if storage != capi.STORE_F32:
            eng.set_storage(storage)
        out.append((spec, rp, sc, eng))
    return out


def _state(eng):
    return (eng.depth(), eng.costs(), eng.used_ids(), eng.rng().view(np.uint32))


def _standalone(engs, waves=1):
    res = []
    for _, _, _, eng in engs:
        eng.set_sweep_waves(waves)
        eng.prepare()
        eng.run()
        res.append(_state(eng))
    return res


def _assert_same(engs, want, tag):
    for i, (_, _, _, eng) in enumerate(engs):
        got = _state(eng)
        for name, a, b in zip(("depth", "costs", "ids", "rng"), got, want[i]):
            assert np.array_equal(a, b), f"{tag}: view {i} {name} differs from the standalone run"


@pytest.fixture(scope="module")
def views(gpu_ctx):
    engs = _engines(gpu_ctx)
    yield engs
    for e in engs:
        e[3].close()


@pytest.mark.parametrize("graph", [0, 1])
def test_batch_equals_standalone(views, gpu_ctx, graph):
    _say(f"case: batch of {len(views)} views vs standalone runs, graph={graph}")
    want = _standalone(views)
    b = capi.Batch(gpu_ctx, [e[3] for e in views])
    b.set_graph(bool(graph))
    _say(f"  launch shape {b.launch()}")
    for rep in range(3):   # replays, with eager work (prepare) between them
        b.prepare()
        b.run()
        _assert_same(views, want, f"replay {rep}")
    b.close()


def test_batch_many_views(gpu_ctx):
    """Members with 3, 20 and 40 source views at 9 draws per row: the batch
    runs the general sweep kernels (the view CDF over the whole wave, draws
    resolved by one compare of floor(64 / V) draws at a time) for every
    member, and each member equals its standalone run."""
    _say("case: batch of views with 3, 20 and 40 sources, S = 9")
    engs = []
    for W, H, V, seed in [(40, 30, 3, 71), (36, 28, 20, 72), (32, 24, 40, 73)]:
        spec, rp = configs.tiny(W=W, H=H, V=V, hw=2, S=9, iters=2, seed=seed)
        sc, eng = make_engine(gpu_ctx, spec, rp, prepare=False)
        engs.append((spec, rp, sc, eng))
    try:
        want = _standalone(engs)
        b = capi.Batch(gpu_ctx, [e[3] for e in engs])
        _say(f"  launch shape {b.launch()}")
        b.prepare()
        b.run()
        _assert_same(engs, want, "many views")
        b.close()
    finally:
        for e in engs:
            e[3].close()


@pytest.mark.parametrize("waves,workgroups", [(1, 8), (2, 0), (2, 24)])
def test_batch_launch_shapes(views, gpu_ctx, waves, workgroups):
    """Few persistent workgroups (each drains many column strips) or two
    waves per strip: same results as the standalone run with those waves."""
    _say(f"case: batch launch shape waves={waves} workgroups={workgroups}")
    want = _standalone(views, waves)
    b = capi.Batch(gpu_ctx, [e[3] for e in views])
    b.set_launch(waves, workgroups)
    b.prepare()
    b.run()
    _assert_same(views, want, f"waves={waves} workgroups={workgroups}")
    b.close()
    for e in views:
        e[3].set_sweep_waves(1)


@pytest.mark.parametrize("hw,storage", [(7, capi.STORE_F32), (7, capi.STORE_F16),
                                        (5, capi.STORE_F32)])
def test_batch_three_waves_per_simd(gpu_ctx, hw, storage):
    """Launch shape 3 (one wave per strip, a register budget of three waves
    per SIMD) at the window sizes it is built for: the same column code as
    one wave per strip, so the same results bit for bit."""
    _say(f"case: batch shape 3, window {2 * hw + 1}, storage {storage}")
    engs = []
    for W, H, V, seed in [(64, 48, 3, 21), (45, 61, 2, 22), (70, 33, 4, 23)]:
        spec, rp = configs.tiny(W=W, H=H, V=V, hw=hw, S=2, iters=2, seed=seed)
        sc, eng = make_engine(gpu_ctx, spec, rp, prepare=False)
        eng.set_storage(storage)
        engs.append((spec, rp, sc, eng))
    want = _standalone(engs, 1)
    b = capi.Batch(gpu_ctx, [e[3] for e in engs])
    for wgs in (0, 24):
        b.set_launch(3, wgs)
        assert b.launch()[0] == 3
        b.prepare()
        b.run()
        _assert_same(engs, want, f"shape 3, {wgs} workgroups")
    b.close()
    for e in engs:
        e[3].close()


def test_batch_fp16(gpu_ctx):
    _say("case: batch with FP16 storage vs standalone FP16 runs")
    engs = _engines(gpu_ctx, capi.STORE_F16, SHAPES[:3])
    want = _standalone(engs)
    b = capi.Batch(gpu_ctx, [e[3] for e in engs])
    b.prepare()
    b.run()
    _assert_same(engs, want, "fp16")
    b.close()
    for e in engs:
        e[3].close()


def test_batch_member_against_oracle(gpu_ctx):
    """The 320x240, 4-source workload batched with a tiny view: >= 99% of its
    final depths within 1% of the FP64 oracle on the same inputs."""
    _say("case: batch member (c1) against the oracle")
    spec, rp = configs.C1_SMALL
    sc, eng = make_engine(gpu_ctx, spec, rp, prepare=False)
    tspec, trp = configs.tiny(W=64, H=48, V=3, hw=5, S=2, iters=3, seed=9)
    tsc, teng = make_engine(gpu_ctx, tspec, trp, prepare=False)
    b = capi.Batch(gpu_ctx, [teng, eng])
    b.prepare()
    d0, st0 = eng.depth(), eng.rng()
    ms = b.run(timed=True)
    dg = eng.depth()
    pb = oracle_for(eng, rp, sc)
    do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples, rp.num_iterations)
    agree = within(dg, do, 0.01)
    _say(f"  {ms:.2f} ms; within 1% of the oracle: {agree:.4f}")
    assert agree >= 0.99
    b.close()
    eng.close()
    teng.close()


def test_batch_errors(gpu_ctx):
    _say("case: batch argument and state errors")
    a = _engines(gpu_ctx, shapes=SHAPES[:1])[0]
    spec, rp = configs.tiny(W=32, H=24, V=2, hw=3, S=2, iters=2)
    sc, other = make_engine(gpu_ctx, spec, rp, prepare=False)
    with pytest.raises(capi.PmxError):
        capi.Batch(gpu_ctx, [a[3], other])          # different half window
    with pytest.raises(capi.PmxError):
        capi.Batch(gpu_ctx, [a[3], a[3]])           # the same engine twice
    b = capi.Batch(gpu_ctx, [a[3]])
    with pytest.raises(capi.PmxError):
        b.run()                                     # not prepared
    b.prepare()
    b.run()
    with pytest.raises(capi.PmxError):
        b.run()                                     # not freshly prepared
    with pytest.raises(capi.PmxError):
        b.set_launch(4, 0)                          # shapes 1, 2 and 3 only
    with pytest.raises(capi.PmxError):
        b.set_launch(0, 0)
    with pytest.raises(capi.PmxError):
        b.set_launch(1, 4)                          # fewer workgroups than XCD queues
    b.close()
    a[3].close()
    other.close()


def test_batch_autotune(views, gpu_ctx):
    _say("case: batch autotuner")
    b = capi.Batch(gpu_ctx, [e[3] for e in views])
    w, g, ms = b.autotune()
    _say(f"  chosen waves={w} workgroups={g}: {ms:.3f} ms")
    assert w in (1, 2, 3) and g >= 8 and ms > 0
    assert b.launch() == (w, g)
    want = _standalone(views, 1 if w == 3 else w)
    b.prepare()
    b.run()
    _assert_same(views, want, "after autotune")
    b.close()
    for e in views:
        e[3].set_sweep_waves(1)


def test_batch_with_proposal(gpu_ctx):
    """A learned proposal installed on one engine of the batch: that view
    equals its standalone run with the same proposal, the others their
    plain runs."""
    _say("case: batch with a proposal on one view")
    engs = _engines(gpu_ctx, shapes=SHAPES[:3])
    spec, rp, sc, e0 = engs[0]
    rng = np.random.default_rng(5)
    prop = (sc.gt_depth * (1 + 0.03 * rng.uniform(-1, 1, sc.gt_depth.shape))).astype(np.float32)
    prop[rng.uniform(size=prop.shape) > 0.7] = 0.0
    spread = np.where(rng.uniform(size=prop.shape) < 0.5, 0.02 / sc.gt_depth, 0.0).astype(np.float32)
    pd = capi.DeviceArray(gpu_ctx, prop.shape, np.float32).upload(prop)
    sd = capi.DeviceArray(gpu_ctx, prop.shape, np.float32).upload(spread)
    e0.set_proposal(pd.ptr, e0.W * 4, sd.ptr)
    want = _standalone(engs)
    b = capi.Batch(gpu_ctx, [e[3] for e in engs])
    for graph in (False, True):
        b.set_graph(graph)
        b.prepare()
        b.run()
        _assert_same(engs, want, f"proposal, graph={graph}")
    b.close()
    e0.set_proposal(None, 0)
    for e in engs:
        e[3].close()
    pd.close()
    sd.close()
