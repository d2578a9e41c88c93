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
        eng.set_sweep_waves(1)      # the batched launch runs one wave per strip
        eng.prepare()
        eng.run()
        want = eng.depth()
        eng.close()
        assert np.array_equal(srv.results[r.rid], want), r
        assert np.array_equal(srv2.results[r.rid], want), r
    ctx2.close()


def test_large_group_runs_shape_3(gpu_ctx):
    """Four 1024x768 requests arriving together form one group with more
    column strips (4 x 768) than launch shape 1 keeps resident, so the group
    runs shape 3; every served depth map still equals a standalone run of
    the same request with one wave per column, bit for bit."""
    _say("case: serving group of four 1024x768 requests (launch shape 3)")
    pool = []
    for i in range(len(serving.SIZES)):
        spec, _ = configs.tiny(W=1024, H=768, V=4, hw=7, S=2, iters=1, seed=60 + i, planes=4)
        pool.append(scene.render(spec))
    _, rp = configs.tiny(hw=7, S=2, iters=1)
    stream = [serving.Request(i, 0.0, i, 3 + i % 2) for i in range(4)]
    srv = serving.Server([gpu_ctx], rp, pool, budget_bytes=2 ** 40)
    srv.keep_results = True
    res = srv.run(stream, log=_say)
    _say(f"  {res['requests']} requests in {res['batches']} group(s), shapes {srv.launch_shapes}")
    assert res["requests"] == 4 and res["batches"] == 1
    assert srv.launch_shapes == [3]
    assert res["groups_by_launch_shape"] == {"3": 1}
    for r in stream:
        sc = pool[r.size]
        eng = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs[:r.n_src]), sc.cams[:r.n_src + 1])
        eng.set_storage(capi.STORE_F16)
        eng.set_sweep_waves(1)
        eng.prepare()
        eng.run()
        want = eng.depth()
        eng.close()
        assert np.array_equal(srv.results[r.rid], want), r


def test_footprint_matches_device_allocation(gpu_ctx):
    """pmx_engine_footprint (the admission measure) against the device memory
    an engine actually takes: hipMemGetInfo before and after creating an
    FP16 engine with a proposal installed."""
    import ctypes

    _say("case: engine footprint vs hipMemGetInfo")
    # the HIP runtime instance libpmx.so already loaded
    path = next(ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln)
    hip = ctypes.CDLL(path)
    free_b, total_b = ctypes.c_size_t(), ctypes.c_size_t()

    def free_now():
        gpu_ctx.sync()
        assert hip.hipMemGetInfo(ctypes.byref(free_b), ctypes.byref(total_b)) == 0
        return free_b.value

    spec, rp = configs.tiny(W=1024, H=768, V=6, hw=7, S=3, iters=1, seed=4)
    sc = scene.render(spec)
    fp = capi.engine_footprint(rp, spec.W, spec.H, [(spec.W, spec.H)] * spec.V,
                               capi.STORE_F16, proposal=True)
    f0 = free_now()
    eng = capi.Engine(gpu_ctx, rp, sc.ref, list(sc.srcs), sc.cams)
    eng.set_storage(capi.STORE_F16)
    prop = capi.DeviceArray(gpu_ctx, (spec.H, spec.W), np.float32)
    eng.set_proposal(prop.ptr, spec.W * 4)
    f2 = free_now()
    used = (f0 - f2) - prop.nbytes
    _say(f"  footprint {fp / 2**20:.1f} MiB, measured {used / 2**20:.1f} MiB")
    # the device allocator rounds each of the ~25 allocations up to its page
    assert 0.97 * fp <= used <= fp + 64 * 2**20
    eng.close()
    prop.close()
