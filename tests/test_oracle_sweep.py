"""End-to-end checks of the oracle's sweeps and schedule (topToDown/downToTop
and runPatchMatch, src/patchMatch.cu:263-407, :818-1209): convergence toward
the rendered ground truth, invariance to the OpenMP thread count, and the
structural quirks (first row untouched, row-0/column-0 generator states,
used-view ids written by the last sweep only)."""
import numpy as np
import pytest

from oracle import pmx_oracle as O
from patchmatch_cuda_amd import configs, scene


def f32(x):
    return float(np.float32(x))


def _setup(W=96, H=72, V=3, hw=3, S=2, iters=3):
    spec, rp = configs.tiny(W=W, H=H, V=V, hw=hw, S=S, iters=iters)
    sc = scene.render(spec)
    ref = O.blur(sc.ref, rp.gaussian_sigma, rp.blur_mode)
    srcs = np.stack([O.blur(s, rp.gaussian_sigma, rp.blur_mode) for s in sc.srcs])
    pb = O.Problem(ref, srcs, O.homographies(sc.cams), hw, f32(rp.depth_near), f32(rp.depth_far),
                   f32(rp.spm_alpha), O.inv_k4(sc.cams[0, :9]))
    st = O.rng_init_grid(W, H, rp.block_x, rp.block_y, rp.rng_seed)
    d0 = O.init_depth(st, f32(rp.depth_near), f32(rp.depth_far), rp.init_mode)
    return sc, rp, pb, st, d0


@pytest.fixture(scope="module")
def setup():
    return _setup()


def test_run_converges_toward_ground_truth(setup):
    sc, rp, pb, st, d0 = setup
    d, cost, ids = pb.run(d0, st.copy(), rp.num_samples, rp.num_iterations)
    gt = sc.gt_depth.astype(np.float64)
    hw = rp.half_window
    m = np.zeros_like(gt, bool)
    m[hw:-hw, hw:-hw] = True
    rel = np.abs(d - gt) / gt
    init_rel = np.abs(d0 - gt) / gt
    assert (init_rel[m] < 0.05).mean() < 0.1
    assert (rel[m] < 0.05).mean() > 0.55
    assert np.median(rel[m]) < 0.05
    assert np.all((d >= f32(rp.depth_near) * 0.999) & (d <= f32(rp.depth_far) * 1.001))
    # away from the left/top border (where the tile-index quirk pairs shifted
    # values, so |NCC| may exceed 1) a cost is 1 - NCC in [0, 2]
    inner = cost[:, hw:, hw:]
    assert np.all((inner >= -1e-9) & (inner <= 2.0 + 1e-9))
    assert np.all(np.isfinite(cost))


def test_thread_count_invariance(setup):
    sc, rp, pb, st, d0 = setup
    old = O.get_threads()
    try:
        O.set_threads(1)
        s1 = st.copy()
        a = pb.run(d0, s1, rp.num_samples, 2)
        O.set_threads(4)
        s4 = st.copy()
        b = pb.run(d0, s4, rp.num_samples, 2)
    finally:
        O.set_threads(old)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(s1, s4)


@pytest.mark.parametrize("rotated", [False, True])
@pytest.mark.parametrize("down", [True, False])
def test_sweep_structure(setup, rotated, down):
    sc, rp, pb, st, d0 = setup
    H, W = d0.shape
    depth_s = np.ascontiguousarray(d0.T if rotated else d0)
    cost_s = pb.cost_all_views(depth_s, rotated=rotated)
    states = st.copy()
    before_depth = depth_s.copy()
    ids, _ = pb.sweep(depth_s, cost_s, states, down, rotated, 2, record_ids=True)
    first = 0 if down else depth_s.shape[0] - 1
    assert np.array_equal(depth_s[first], before_depth[first])          # first row never updated
    assert np.all(ids[:, first] == 0)                                   # nor are its view ids
    assert np.any(ids[:, 1:-1] != 0)
    assert ids.max() < pb.V
    if rotated:                                                         # column-0 states only
        assert np.array_equal(states[:, 1:], st[:, 1:])
        assert not np.array_equal(states[:, 0], st[:, 0])
    else:                                                               # row-0 states only
        assert np.array_equal(states[1:], st[1:])
        assert not np.array_equal(states[0], st[0])
    assert np.mean(depth_s != before_depth) > 0.2


def test_cost_all_views_counts(setup):
    sc, rp, pb, st, d0 = setup
    e0, s0 = pb.counters()
    pb.cost_all_views(np.ascontiguousarray(d0.T), rotated=True)
    e1, s1 = pb.counters()
    H, W = d0.shape
    hw = rp.half_window
    span = lambda n: sum(min(i + hw, n - 1) - max(i - hw, 0) + 1 for i in range(n))
    assert e1 - e0 == W * H * pb.V
    assert s1 - s0 == span(W) * span(H) * pb.V


def _mix32(v):
    v = np.uint64(v) & np.uint64(0xFFFFFFFF)
    v ^= v >> np.uint64(16)
    v = (v * np.uint64(0x7feb352d)) & np.uint64(0xFFFFFFFF)
    v ^= v >> np.uint64(15)
    v = (v * np.uint64(0x846ca68b)) & np.uint64(0xFFFFFFFF)
    v ^= v >> np.uint64(16)
    return v


def test_proposal_draw_restated():
    """The proposal hypothesis: an independent numpy restatement of the hash
    and the FP32 arithmetic, spread 0 = the depth itself, draws inside
    [near, far], and h spread evenly over [-1, 1)."""
    W, near, far = 37, np.float32(1.0), np.float32(20.0)
    f = np.float32
    hs = []
    for k in range(3):
        for y in range(5):
            for x in range(W):
                depth, spread = f(2.0 + 0.1 * x), f(0.05)
                key = _mix32(np.uint64(y * W + x) ^ _mix32(
                    (np.uint64(k) * np.uint64(0x9E3779B9) + np.uint64(0x632BE5AB)) & np.uint64(0xFFFFFFFF)))
                h = f(int(key >> np.uint64(8))) * f(1.0 / 8388608.0) - f(1.0)
                # fmaf: the product of two floats is exact in double, and the
                # sum is rounded once more to float
                u = f(float(spread) * float(h) + float(f(1.0) / depth))
                u = min(max(u, f(1.0) / far), f(1.0) / near)
                want = float(f(1.0) / f(u))
                got = O.proposal_draw(depth, spread, x, y, W, float(near), float(far), k)
                assert got == want, (k, x, y, got, want)
                assert float(near) <= got <= float(far)
                hs.append(float(h))
                assert O.proposal_draw(depth, 0.0, x, y, W, 1.0, 20.0, k) == float(depth)
    hs = np.array(hs)
    assert hs.min() >= -1.0 and hs.max() < 1.0
    assert abs(hs.mean()) < 0.1 and np.histogram(hs, 4, (-1, 1))[0].min() > len(hs) / 8


def test_proposal_sweeps_limit(setup):
    """pmxo_problem_set_proposal_sweeps: a limit covering every sweep of the
    run (or 0) leaves the run equal, bit for bit, to one with the proposal in
    every sweep; a limit of one iteration's 4 sweeps changes a 3-iteration
    run but not a 1-iteration one."""
    sc, rp, pb, st, d0 = setup
    prop = (sc.gt_depth * 1.02).astype(np.float64)
    spread = np.full(prop.shape, 0.02)

    def run(limit, iters):
        pb.set_proposal(prop, spread)
        pb.set_proposal_sweeps(limit)
        try:
            return pb.run(d0, st.copy(), rp.num_samples, iters)[0]
        finally:
            pb.set_proposal_sweeps(0)
            pb.set_proposal(None)

    every = run(0, 3)
    assert np.array_equal(run(12, 3), every)
    assert np.array_equal(run(1000, 3), every)
    assert np.array_equal(run(4, 1), run(0, 1))
    first = run(4, 3)
    assert not np.array_equal(first, every)
    plain = pb.run(d0, st.copy(), rp.num_samples, 3)[0]
    assert not np.array_equal(first, plain)


def test_trace_lpost_max(setup):
    """The traced log of the largest posterior numerator (m_f * beta of the
    row's views): the log of a probability product on every processed row,
    +inf on the first row, and a finite, normalised CDF wherever it is far
    from FP64 underflow."""
    sc, rp, pb, st, d0 = setup
    depth_s = np.ascontiguousarray(d0)
    cost_s = pb.cost_all_views(depth_s, rotated=False)
    _, tr = pb.sweep(depth_s, cost_s, st.copy(), 1, 0, 2, trace="full")
    lp = tr["lpost_max"]
    assert np.all(np.isposinf(lp[0]))
    body = lp[1:]
    assert np.all(body <= 0.0)
    ok = body > -700.0
    assert ok.mean() > 0.9
    np.testing.assert_allclose(tr["cdf"][1:][ok][:, -1], 1.0, rtol=0, atol=1e-12)
