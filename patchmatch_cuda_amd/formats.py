"""File formats of the reference program, for preparing inputs to the
command-line tool (pmx_cli) and reading its outputs.

* Middlebury camera lists (src/image.cpp:64-91 of the reference):
  ``N`` then N lines ``name k11..k33 r11..r33 t1 t2 t3`` with T = -R C.
* NVM_V3 models (src/image.cpp:120-244): cameras as focal length,
  quaternion (w, x, y, z), centre and two distortion terms, then 3D points
  with their observations.
* The INI parameter file of src/main.cpp:42-91 ([params] section).
* Depth-map files of Array2D_wrapper::saveToFile
  (src/array2d_wrapper.h:51-70): ASCII ``W&H&D&`` then W*H*D raw values.
* PGM/PNG writers for 8-bit gray images.

The C readers of the same formats live in libpmx_drop_in.so (cpp/pmx_io.cpp);
``io_lib()`` binds their test entry points.
"""
from __future__ import annotations

import ctypes
import os
import struct
import zlib

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DROPIN_LIB = os.path.join(_HERE, "libpmx_drop_in.so")
CLI = os.path.join(_HERE, "pmx_cli")


# ------------------------------------------------------------ images
def write_pgm(path, img):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(f"P5\n{w} {h}\n255\n".encode())
        f.write(img.tobytes())


def write_png(path, img, filt=None, bit_depth=8, palette=None):
    """Gray (H, W), gray+alpha (H, W, 2), RGB (H, W, 3) or RGBA (H, W, 4)
    image; ``filt`` picks one PNG filter type for every row (default: row
    index mod 5, so all five filters occur); ``palette`` (k x 3 uint8) writes
    an indexed image of the given indices."""
    a = np.asarray(img)
    if palette is not None:
        ctype, ch = 3, 1
        a = a.astype(np.uint8)
    else:
        ch = 1 if a.ndim == 2 else a.shape[2]
        ctype = {1: 0, 2: 4, 3: 2, 4: 6}[ch]
    h, w = a.shape[:2]
    if bit_depth == 16:
        raw = a.astype(">u2").reshape(h, w * ch).view(np.uint8).reshape(h, w * ch * 2)
    else:
        raw = a.astype(np.uint8).reshape(h, w * ch)
    bpp = ch * bit_depth // 8
    rows = []
    prev = np.zeros(raw.shape[1], np.int32)
    for y in range(h):
        cur = raw[y].astype(np.int32)
        ft = (y % 5) if filt is None else filt
        left = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])
        upleft = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        if ft == 0:
            out = cur
        elif ft == 1:
            out = cur - left
        elif ft == 2:
            out = cur - prev
        elif ft == 3:
            out = cur - ((left + prev) >> 1)
        else:
            p = left + prev - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
            out = cur - pred
        rows.append(bytes([ft]) + (out & 0xFF).astype(np.uint8).tobytes())
        prev = cur

    def chunk(t, d):
        c = struct.pack(">I", len(d)) + t + d
        return c + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    ihdr = struct.pack(">IIBBBBB", w, h, bit_depth, ctype, 0, 0, 0)
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr)
    if palette is not None:
        data += chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    data += chunk(b"IDAT", zlib.compress(b"".join(rows))) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def gray_bgr2gray(rgb):
    """OpenCV's 8-bit BGR2GRAY weights (fixed point, 14 fraction bits)."""
    rgb = np.asarray(rgb, np.uint32)
    return ((rgb[..., 0] * 4899 + rgb[..., 1] * 9617 + rgb[..., 2] * 1868 + 8192) >> 14).astype(np.uint8)


# ----------------------------------------------------------- cameras
def cams_to_krt(cams):
    """(V+1) x 21 rows (K, R, C) -> lists of K, R, T = -R C (float64)."""
    cams = np.asarray(cams, np.float64)
    K = cams[:, :9].reshape(-1, 3, 3)
    R = cams[:, 9:18].reshape(-1, 3, 3)
    C = cams[:, 18:21]
    T = -np.einsum("nij,nj->ni", R, C)
    return K, R, T


def write_middlebury(path, names, K, R, T):
    with open(path, "w") as f:
        f.write(f"{len(names)}\n")
        for n, k, r, t in zip(names, K, R, T):
            vals = list(np.asarray(k).ravel()) + list(np.asarray(r).ravel()) + list(np.asarray(t).ravel())
            f.write(n + " " + " ".join(repr(float(v)) for v in vals) + "\n")


def rotation_to_quaternion(R):
    """Unit quaternion (w, x, y, z) of a rotation matrix (Shepperd's method)."""
    R = np.asarray(R, np.float64)
    tr = np.trace(R)
    if tr > 0:
        s = np.sqrt(tr + 1.0) * 2
        q = [0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s]
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        q = [(R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s]
    elif R[1, 1] > R[2, 2]:
        s = np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        q = [(R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s]
    else:
        s = np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        q = [(R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s]
    return np.array(q)


def write_nvm(path, names, focals, Rs, Cs, points, observations):
    """points: P x 3; observations: per point a list of camera indices."""
    with open(path, "w") as f:
        f.write("NVM_V3\n\n")
        f.write(f"{len(names)}\n")
        for n, fo, R, C in zip(names, focals, Rs, Cs):
            q = rotation_to_quaternion(R)
            f.write(f"{n} {fo!r} " + " ".join(repr(float(v)) for v in q) + " "
                    + " ".join(repr(float(v)) for v in C) + " 0 0\n")
        f.write(f"\n{len(points)}\n")
        for p, obs in zip(points, observations):
            f.write(" ".join(repr(float(v)) for v in p) + " 128 128 128 " + str(len(obs)))
            for j, c in enumerate(obs):
                f.write(f" {c} {j} 0.0 0.0")
            f.write("\n")


def write_ini(path, params: dict):
    with open(path, "w") as f:
        f.write("[params]\n")
        for k, v in params.items():
            f.write(f"{k} = {v}\n")


def read_depth_file(path):
    """Array2D_wrapper::saveToFile output -> (depth, height, width) float32 array."""
    with open(path, "rb") as f:
        data = f.read()
    head = []
    pos = 0
    for _ in range(3):
        amp = data.index(b"&", pos)
        head.append(int(data[pos:amp]))
        pos = amp + 1
    w, h, d = head
    arr = np.frombuffer(data, np.float32, count=w * h * d, offset=pos)
    return arr.reshape(d, h, w).copy()


# ------------------------------------------------- C readers (test hooks)
_io = None


def io_lib():
    global _io
    if _io is None:
        if not os.path.exists(DROPIN_LIB):
            raise RuntimeError(f"{DROPIN_LIB} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(DROPIN_LIB)
        i, vp, cp = ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p
        pi = ctypes.POINTER(i)
        L.pmx_io_read_gray.argtypes = [cp, pi, pi, vp, ctypes.c_long]
        L.pmx_io_read_cameras.argtypes = [cp, pi, vp, vp, vp, vp, vp, vp, vp, i]
        L.pmx_io_read_depth_file.argtypes = [cp, pi, pi, pi, vp, ctypes.c_long]
        _io = L
    return _io


def c_read_gray(path):
    L = io_lib()
    w, h = ctypes.c_int(), ctypes.c_int()
    rc = L.pmx_io_read_gray(path.encode(), ctypes.byref(w), ctypes.byref(h), None, 0)
    if rc:
        raise ValueError(f"cannot decode {path}")
    out = np.zeros((h.value, w.value), np.uint8)
    rc = L.pmx_io_read_gray(path.encode(), ctypes.byref(w), ctypes.byref(h), out.ctypes.data, out.size)
    if rc:
        raise ValueError(f"cannot decode {path}")
    return out


def c_read_cameras(path, cap=64):
    """-> (kind, dict of K, R, T, C, widths, heights, near_far) as the C++ reader sees them."""
    L = io_lib()
    n = ctypes.c_int()
    K = np.zeros((cap, 9), np.float32)
    R = np.zeros((cap, 9), np.float32)
    T = np.zeros((cap, 3), np.float32)
    C = np.zeros((cap, 3), np.float32)
    ws = np.zeros(cap, np.int32)
    hs = np.zeros(cap, np.int32)
    nf = np.zeros((cap, 2), np.float32)
    kind = L.pmx_io_read_cameras(path.encode(), ctypes.byref(n), K.ctypes.data, R.ctypes.data,
                                 T.ctypes.data, C.ctypes.data, ws.ctypes.data, hs.ctypes.data,
                                 nf.ctypes.data, cap)
    m = n.value
    return kind, dict(K=K[:m], R=R[:m], T=T[:m], C=C[:m], widths=ws[:m], heights=hs[:m],
                      near_far=nf[:m])
