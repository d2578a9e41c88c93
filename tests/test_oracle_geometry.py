"""Pins the oracle's homography composition (src/image.cpp:4-13) and the
plane-induced warp T = H1 - H2/(d*scale) (src/patchMatch.cu:465-473) against
projecting 3D points through the two cameras."""
import numpy as np
from scipy.spatial.transform import Rotation

from oracle import pmx_oracle as O


def _cams(seed):
    rs = np.random.default_rng(seed)
    Kr = np.array([[300.0, 0, 160.5], [0, 310.0, 119.0], [0, 0, 1]])
    Ks = np.array([[290.0, 0, 150.0], [0, 305.0, 125.5], [0, 0, 1]])
    Rr = Rotation.from_rotvec(rs.normal(0, 0.1, 3)).as_matrix()
    Rs = Rotation.from_rotvec(rs.normal(0, 0.1, 3)).as_matrix()
    Cr = rs.normal(0, 0.3, 3)
    Cs = Cr + rs.normal(0, 0.4, 3)
    return Kr, Rr, Cr, Ks, Rs, Cs


def _project(K, R, C, Xw):
    x = K @ (R @ (Xw - C))
    return x[:2] / x[2]


def test_identity_cameras():
    K = np.array([[250.0, 0, 100], [0, 250, 80], [0, 0, 1]])
    H1, H2 = O.compose(K, np.eye(3), np.zeros(3), K, np.eye(3), np.zeros(3))
    np.testing.assert_allclose(H1, np.eye(3), atol=1e-13)
    assert np.all(H2 == 0)


def test_fronto_parallel_plane_points_map_through_both_cameras():
    for seed in range(5):
        Kr, Rr, Cr, Ks, Rs, Cs = _cams(seed)
        H1, H2 = O.compose(Kr, Rr, Cr, Ks, Rs, Cs, 0.0, 1.0)
        for (x, y, d) in [(10.5, 20.5, 3.0), (160.5, 120.5, 7.5), (300.5, 5.5, 1.2)]:
            p = np.array([x, y, 1.0])
            Xc = d * np.linalg.solve(Kr, p)              # z-depth d on the ray
            Xw = Rr.T @ Xc + Cr
            T = H1 - H2 / d                               # scale == 1 for n = (0,0,1)
            q = T @ p
            np.testing.assert_allclose(q[:2] / q[2], _project(Ks, Rs, Cs, Xw), rtol=1e-11, atol=1e-9)


def test_slanted_plane_window_points():
    """With n = (oX, 0, oZ) the window around p lies on the plane through the
    centre point at z-depth d: n . X = d * scale(p), scale(p) = n . Kr^-1 p."""
    oX, oZ = 0.6, 0.8
    for seed in range(5):
        Kr, Rr, Cr, Ks, Rs, Cs = _cams(10 + seed)
        H1, H2 = O.compose(Kr, Rr, Cr, Ks, Rs, Cs, oX, oZ)
        n = np.array([oX, 0.0, oZ])
        Kri = np.linalg.inv(Kr)
        p = np.array([140.5, 90.5, 1.0])
        d = 4.0
        scale = n @ (Kri @ p)
        T = H1 - H2 / (d * scale)
        for dx, dy in [(0, 0), (-7, 3), (7, -7), (5, 6)]:
            q = p + np.array([dx, dy, 0.0])
            mu = d * scale / (n @ (Kri @ q))
            Xw = Rr.T @ (mu * (Kri @ q)) + Cr
            t = T @ q
            np.testing.assert_allclose(t[:2] / t[2], _project(Ks, Rs, Cs, Xw), rtol=1e-11, atol=1e-9)


def test_homographies_rows_match_compose():
    Kr, Rr, Cr, Ks, Rs, Cs = _cams(3)
    cams = np.zeros((2, 21))
    cams[0] = np.concatenate([Kr.ravel(), Rr.ravel(), Cr])
    cams[1] = np.concatenate([Ks.ravel(), Rs.ravel(), Cs])
    Hm = O.homographies(cams)
    H1, H2 = O.compose(Kr, Rr, Cr, Ks, Rs, Cs)
    assert np.array_equal(Hm[0, :9], H1.ravel()) and np.array_equal(Hm[0, 9:], H2.ravel())
