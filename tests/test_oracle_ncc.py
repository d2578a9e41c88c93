"""Pins the oracle's matching cost 1 - NCC (computeNCC + readImageIntoSharedMemory,
src/patchMatch.cu:433-667) by closed forms, by an independent brute force
whose bilinear sampling comes from scipy, and by the left-border quirk of the
reference's tile index (src/patchMatch.cu:536-551)."""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

from oracle import pmx_oracle as O

H, W, HW = 20, 24, 2
SHIFT = np.array([[1.0, 0, -0.5], [0, 1, -0.5], [0, 0, 1]])   # maps pixel centre (c+.5, r+.5) to texel (c, r)


def _problem(ref, srcs, H1s, H2s=None, hw=HW, oX=0.0, oZ=1.0, invK=(1 / 300.0, -0.4, 1 / 300.0, -0.3)):
    V = len(srcs)
    Hm = np.zeros((V, 18))
    for v in range(V):
        Hm[v, :9] = np.asarray(H1s[v]).ravel()
        if H2s is not None:
            Hm[v, 9:] = np.asarray(H2s[v]).ravel()
    return O.Problem(ref, np.asarray(srcs, np.float64), Hm, hw, 1.0, 10.0, 0.45, list(invK), oX, oZ)


@pytest.fixture(scope="module")
def ref():
    return np.random.default_rng(1).random((H, W))


def test_identical_patch_costs_zero(ref):
    pb = _problem(ref, [ref], [SHIFT])
    for r in [0, 1, 7, H - 1]:
        for c in [HW, 9, W - 2, W - 1]:
            assert abs(pb.ncc(0, r, c, 3.0)) < 1e-12, (r, c)


@pytest.mark.parametrize("a,b,expect", [(2.5, 0.3, 0.0), (0.4, -1.0, 0.0), (-1.7, 3.0, 2.0)])
def test_affine_intensity(ref, a, b, expect):
    pb = _problem(ref, [a * ref + b], [SHIFT])
    for (r, c) in [(5, 5), (0, 10), (H - 1, W - 1), (10, HW)]:
        assert abs(pb.ncc(0, r, c, 5.0) - expect) < 1e-11


def test_constant_source_costs_one(ref):
    pb = _problem(ref, [np.full((H, W), 0.6)], [SHIFT])
    assert pb.ncc(0, 8, 8, 2.0) == 1.0
    # constant reference window as well: the reference variance is clamped too
    pb2 = _problem(np.full((H, W), 0.2), [ref], [SHIFT])
    assert pb2.ncc(0, 8, 8, 2.0) == 1.0


def test_out_of_frame_is_sentinel(ref):
    far = SHIFT.copy()
    far[0, 2] += 1000.0
    pb = _problem(ref, [ref], [far])
    assert pb.ncc(0, 8, 8, 2.0) == 2.0
    assert pb.ncc(0, 0, 0, 2.0) == 2.0


def _ncc_formula(ip, rv, sI, sII):
    n = len(ip)
    c1 = sII - sI * sI / n
    c2 = (ip * ip).sum() - ip.sum() ** 2 / n
    c1 = 0.0 if c1 < 1e-5 else c1
    c2 = 0.0 if c2 < 1e-5 else c2
    den = np.sqrt(c1 * c2)
    if (ip * ip).sum() == 0.0:
        return 2.0
    if den == 0.0:
        return 1.0
    return 1.0 - ((ip * rv).sum() - sI * ip.sum() / n) / den


def test_left_border_quirk(ref):
    """Columns c < hw pair in-bounds source samples with reference values
    shifted left by hw - c (zeros from the border)."""
    pb = _problem(ref, [ref], [SHIFT])
    sI, sII = O.box_sums(ref, HW)
    r = 9
    for c in range(HW):
        ip, rv = [], []
        for rr in range(r - HW, r + HW + 1):
            j = 0
            for cc in range(c - HW, c + HW + 1):
                if cc < 0:
                    continue
                ip.append(ref[rr, cc])
                col = c - HW + j
                rv.append(ref[rr, col] if col >= 0 else 0.0)
                j += 1
        expect = _ncc_formula(np.array(ip), np.array(rv), sI[r, c], sII[r, c])
        got = pb.ncc(0, r, c, 3.0)
        assert abs(got - expect) < 1e-12
        assert got > 0.05          # identical images, yet not a perfect match


def test_projective_brute_force_with_scipy_bilinear():
    rs = np.random.default_rng(7)
    ref = rs.random((H, W))
    srcs = [rs.random((H + 3, W + 5)), rs.random((H + 3, W + 5))]
    H1s, H2s = [], []
    for v in range(2):
        H1 = np.eye(3) + rs.normal(0, 0.02, (3, 3))
        H1[2, :2] = rs.normal(0, 2e-4, 2)
        H1[:2, 2] += rs.normal(0, 1.0, 2)
        H1s.append(H1)
        H2s.append(rs.normal(0, 0.5, (3, 3)) * np.array([[1, 1, 1], [1, 1, 1], [1e-3, 1e-3, 1e-2]]))
    invK = (1 / 280.0, -0.45, 1 / 280.0, -0.35)
    oX, oZ = 0.3, 0.95
    pb = _problem(ref, srcs, H1s, H2s, hw=3, oX=oX, oZ=oZ, invK=invK)
    sI, sII = O.box_sums(ref, 3)
    for v in range(2):
        for (r, c, d) in [(5, 5, 2.0), (10, 13, 4.5), (H - 1, W - 1, 7.0), (0, 20, 1.5)]:
            scale = oX * (invK[0] * (c + 0.5) + invK[1]) + oZ
            T = H1s[v] - H2s[v] / d / scale
            ip, rv = [], []
            for rr in range(r - 3, r + 4):
                for cc in range(c - 3, c + 4):
                    if not (0 <= rr < H and 0 <= cc < W):
                        continue
                    q = T @ np.array([cc + 0.5, rr + 0.5, 1.0])
                    ip.append(map_coordinates(srcs[v], [[q[1] / q[2]], [q[0] / q[2]]], order=1,
                                              mode="grid-constant", cval=0.0)[0])
                    rv.append(ref[rr, cc])
            expect = _ncc_formula(np.array(ip), np.array(rv), sI[r, c], sII[r, c])
            assert abs(pb.ncc(v, r, c, d) - expect) < 1e-10, (v, r, c)


def test_rotated_frame_equals_plain_frame_in_interior(ref):
    rs = np.random.default_rng(3)
    src = rs.random((H, W))
    H1 = SHIFT.copy()
    H1[0, 1] = 0.05
    H2 = np.zeros((3, 3))
    H2[0, 2] = 1.5
    pb = _problem(ref, [src], [H1], [H2])
    for (x, y) in [(HW, HW), (9, 4), (W - 1, H - 1), (15, 11)]:
        a = pb.ncc(0, y, x, 3.3, rotated=False)
        b = pb.ncc(0, x, y, 3.3, rotated=True)
        assert abs(a - b) < 1e-12


def test_cost_all_views_matches_pointwise(ref):
    rs = np.random.default_rng(5)
    srcs = [ref, 0.5 * ref + 0.1, rs.random((H, W))]
    pb = _problem(ref, srcs, [SHIFT] * 3)
    depth_s = rs.uniform(2, 8, (W, H))              # rotated frame: Hs = W, Ws = H
    cost = pb.cost_all_views(depth_s, rotated=True)
    for v in range(3):
        for (rr, cc) in [(3, 4), (W - 1, 0), (10, H - 1)]:
            assert cost[v, rr, cc] == pb.ncc(v, rr, cc, depth_s[rr, cc], rotated=True)
