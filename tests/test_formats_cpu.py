"""File formats of the reference program (cpp/pmx_io.cpp, formats.py): image
decoding, Middlebury and NVM camera lists, depth files, and the command-line
tool's configuration errors.  CPU only."""
import os
import subprocess

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from patchmatch_cuda_amd import formats as F


@pytest.fixture
def tmp(tmp_path):
    return str(tmp_path)


def test_png_all_filters_gray_and_color(tmp):
    rs = np.random.default_rng(0)
    g = rs.integers(0, 256, (37, 29), dtype=np.uint8)
    for filt in [None, 0, 1, 2, 3, 4]:
        p = os.path.join(tmp, f"g{filt}.png")
        F.write_png(p, g, filt=filt)
        assert np.array_equal(F.c_read_gray(p), g)
    rgb = rs.integers(0, 256, (21, 17, 3), dtype=np.uint8)
    p = os.path.join(tmp, "rgb.png")
    F.write_png(p, rgb)
    assert np.array_equal(F.c_read_gray(p), F.gray_bgr2gray(rgb))
    rgba = rs.integers(0, 256, (9, 14, 4), dtype=np.uint8)
    F.write_png(p, rgba)
    assert np.array_equal(F.c_read_gray(p), F.gray_bgr2gray(rgba[..., :3]))
    ga = rs.integers(0, 256, (9, 14, 2), dtype=np.uint8)
    F.write_png(p, ga)
    assert np.array_equal(F.c_read_gray(p), ga[..., 0])


def test_png_16bit_and_palette(tmp):
    rs = np.random.default_rng(1)
    g16 = rs.integers(0, 65536, (11, 13), dtype=np.uint16)
    p = os.path.join(tmp, "g16.png")
    F.write_png(p, g16, bit_depth=16)
    assert np.array_equal(F.c_read_gray(p), (g16 >> 8).astype(np.uint8))
    pal = rs.integers(0, 256, (7, 3), dtype=np.uint8)
    idx = rs.integers(0, 7, (10, 12), dtype=np.uint8)
    F.write_png(p, idx, palette=pal)
    assert np.array_equal(F.c_read_gray(p), F.gray_bgr2gray(pal[idx]))


def test_pgm_binary_and_ascii(tmp):
    g = np.random.default_rng(2).integers(0, 256, (8, 5), dtype=np.uint8)
    p = os.path.join(tmp, "a.pgm")
    F.write_pgm(p, g)
    assert np.array_equal(F.c_read_gray(p), g)
    with open(p, "w") as f:
        f.write("P2\n# comment\n5 8\n255\n" + " ".join(str(v) for v in g.ravel()) + "\n")
    assert np.array_equal(F.c_read_gray(p), g)
    with open(p, "wb") as f:
        f.write(b"P7\n1 1\n255\n\x00")
    with pytest.raises(ValueError):
        F.c_read_gray(p)


def _scene_cams(n, seed):
    rs = np.random.default_rng(seed)
    K = np.array([[300.0, 0, 40.0], [0, 300.0, 30.0], [0, 0, 1]])
    Rs = [Rotation.from_rotvec(rs.normal(0, 0.1, 3)).as_matrix() for _ in range(n)]
    Cs = [rs.normal(0, 0.5, 3) for _ in range(n)]
    return K, Rs, Cs


def test_middlebury_round_trip(tmp):
    K, Rs, Cs = _scene_cams(3, 3)
    names = []
    for i in range(3):
        names.append(f"im{i}.pgm")
        F.write_pgm(os.path.join(tmp, names[-1]), np.full((60, 80), 10 * i, np.uint8))
    T = [-R @ C for R, C in zip(Rs, Cs)]
    par = os.path.join(tmp, "cams_par.txt")
    F.write_middlebury(par, names, [K] * 3, Rs, T)
    kind, c = F.c_read_cameras(par)
    assert kind == 2 and len(c["K"]) == 3
    for i in range(3):
        np.testing.assert_allclose(c["K"][i], K.ravel(), rtol=1e-7)
        np.testing.assert_allclose(c["R"][i], Rs[i].ravel(), atol=1e-7)
        np.testing.assert_allclose(c["C"][i], Cs[i], atol=1e-6)   # C = -R^T T
    assert list(c["widths"]) == [80] * 3 and list(c["heights"]) == [60] * 3
    # a missing image fails the whole list, as a failed imread would
    os.remove(os.path.join(tmp, names[1]))
    assert F.c_read_cameras(par)[0] == 0


def test_nvm_depth_ranges(tmp):
    K, Rs, Cs = _scene_cams(2, 4)
    names = ["a.pgm", "b.pgm"]
    for n in names:
        F.write_pgm(os.path.join(tmp, n), np.zeros((60, 80), np.uint8))
    rs = np.random.default_rng(5)
    pts = rs.normal(0, 0.5, (40, 3)) + np.array([0, 0, 6.0])
    obs = [[0, 1] if i % 3 else [1] for i in range(40)]
    nvm = os.path.join(tmp, "model.nvm")
    F.write_nvm(nvm, names, [300.0, 310.0], Rs, Cs, pts, obs)
    kind, c = F.c_read_cameras(nvm)
    assert kind == 1
    for i in range(2):
        np.testing.assert_allclose(c["R"][i], Rs[i].ravel(), atol=1e-6)
        np.testing.assert_allclose(c["C"][i], Cs[i], atol=1e-5)
        assert c["K"][i][2] == 40.0 and c["K"][i][5] == 30.0     # principal point at the centre
        z = np.array([(Rs[i] @ (p - Cs[i]))[2] for p, o in zip(pts, obs) if i in o])
        lo, hi = z.min(), z.max()
        size = hi - lo
        ratio = np.float32(0.25)
        near = lo - ratio * size
        while near <= 0:
            ratio /= np.float32(1.1)
            near = lo - ratio * size
        np.testing.assert_allclose(c["near_far"][i], [near, hi + 0.25 * size], rtol=1e-5)


def test_depth_file_reader(tmp):
    d = np.random.default_rng(6).random((2, 5, 7)).astype(np.float32)
    p = os.path.join(tmp, "depth.bin")
    with open(p, "wb") as f:
        f.write(b"7&5&2&")
        f.write(d.tobytes())
    assert np.array_equal(F.read_depth_file(p), d)


def test_cli_argument_and_config_errors(tmp):
    r = subprocess.run([F.CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Initialization file is mandatory" in r.stdout
    ini = os.path.join(tmp, "c.ini")
    F.write_ini(ini, {"imageInfoFileName": "x.txt", "refImageId": 0})
    r = subprocess.run([F.CLI, ini], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Problems with config file" in r.stdout
    assert "halfWindowSize" in r.stdout
