"""GPU test of the serving loop on a short stream of small requests: every
request completes, a tight HBM budget splits the stream into several groups
on the two alternating contexts, and each served depth map equals a
standalone run of the same request bit for bit."""
import sys

import numpy as np
import pytest

from patchmatch_cuda_amd import capi, configs, scene, serving

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_served_depths_equal_standalone_runs(gpu_ctx):
    _say("case: serving loop, 5 small scene sizes, budget of 3 requests")
    sizes = [(40, 30), (48, 36), (64, 48), (32, 40), (56, 42)]
    pool = []
    for i, (W, H) in enumerate(sizes):
        spec, _ = configs.tiny(W=W, H=H, V=4, hw=2, S=2, iters=2, seed=20 + i)
        pool.append(scene.render(spec))
    _, rp = configs.tiny(hw=2, S=2, iters=2)
    stream = serving.make_stream(20.0, 1.0, seed=3, min_src=2, max_src=4)
    assert len(stream) >= 8
    ctx2 = capi.Context(0)
    srv = serving.Server([gpu_ctx, ctx2], rp, pool, budget_bytes=1)
    # a budget below one request: every group holds exactly one request
    srv.keep_results = True
    res = srv.run(stream, log=_say)
    _say(f"  {res['requests']} requests in {res['batches']} groups, "
         f"p50 {res['latency_s']['p50'] * 1e3:.1f} ms")
    assert res["requests"] == len(stream)
    assert res["batches"] == len(stream)
    assert set(srv.results) == {r.rid for r in stream}
    assert srv.resident == 0
    # a looser budget groups requests
    srv2 = serving.Server([gpu_ctx, ctx2], rp, pool, budget_bytes=2 ** 40)
    srv2.keep_results = True
    res2 = srv2.run(stream, log=_say)
    assert res2["requests"] == len(stream)
    for r in stream[:6]:
        sc = pool[r.size]
        eng = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs[:r.n_src]), sc.cams[:r.n_src + 1])
        eng.set_storage(capi.STORE_F16)
        eng.prepare()
        eng.run()
        want = eng.depth()
        eng.close()
        assert np.array_equal(srv.results[r.rid], want), r
        assert np.array_equal(srv2.results[r.rid], want), r
    ctx2.close()
