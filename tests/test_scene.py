"""The seeded synthetic renderer that stands in for the reference's NVM /
Middlebury inputs: deterministic, multi-view consistent, depths in range."""
import numpy as np

from oracle import pmx_oracle as O
from patchmatch_cuda_amd import configs, scene


def test_render_is_deterministic_and_in_range():
    spec, _ = configs.tiny(W=64, H=48, V=3)
    a = scene.render(spec)
    b = scene.render(spec)
    assert np.array_equal(a.ref, b.ref) and np.array_equal(a.srcs, b.srcs)
    assert np.array_equal(a.cams, b.cams) and np.array_equal(a.gt_depth, b.gt_depth)
    assert a.srcs.shape == (3, 48, 64) and a.cams.shape == (4, 21)
    assert a.gt_depth.min() >= spec.depth_min * 0.99 and a.gt_depth.max() <= spec.depth_max * 1.01
    for R in a.R:
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)


def test_ground_truth_depth_is_photo_consistent():
    """Warping a source view to the reference through the ground-truth depth
    reproduces the reference intensities (up to noise and occlusions)."""
    spec, _ = configs.tiny(W=96, H=72, V=2)
    sc = scene.render(spec)
    Hm = O.homographies(sc.cams)
    ys, xs = np.mgrid[8:64:3, 8:88:3]
    err = []
    for v in range(2):
        H1, H2 = Hm[v, :9].reshape(3, 3), Hm[v, 9:].reshape(3, 3)
        for y, x in zip(ys.ravel(), xs.ravel()):
            d = float(sc.gt_depth[y, x])
            q = (H1 - H2 / d) @ np.array([x + 0.5, y + 0.5, 1.0])
            u, w = q[0] / q[2] - 0.5, q[1] / q[2] - 0.5
            iu, iw = int(round(u)), int(round(w))
            if 0 <= iu < 96 and 0 <= iw < 72:
                err.append(abs(float(sc.srcs[v][iw, iu]) - float(sc.ref[y, x])))
    err = np.array(err)
    assert len(err) > 500
    assert np.median(err) < 0.05


def test_crop_reference_moves_principal_point():
    spec, _ = configs.tiny(W=64, H=48, V=2)
    sc = scene.render(spec)
    ref, srcs, cams, gt = scene.crop_reference(sc, 10, 6, 32, 24)
    assert ref.shape == (24, 32) and gt.shape == (24, 32)
    assert np.array_equal(ref, sc.ref[6:30, 10:42])
    assert np.array_equal(gt, sc.gt_depth[6:30, 10:42])
    assert cams[0, 2] == sc.cams[0, 2] - 10 and cams[0, 5] == sc.cams[0, 5] - 6
    assert np.array_equal(srcs, sc.srcs)
    # a cropped pixel keeps its ray: same homography image as the full-frame pixel
    full = O.homographies(sc.cams)[0]
    crop = O.homographies(cams)[0]
    p_full = np.array([10 + 3.5, 6 + 2.5, 1.0])
    p_crop = np.array([3.5, 2.5, 1.0])
    for k in (0, 9):
        a = full[k:k + 9].reshape(3, 3) @ p_full
        b = crop[k:k + 9].reshape(3, 3) @ p_crop
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
