"""FP64 numpy restatement of the proposal network (pmx_cnn.hip / pmx_net.hip).

TEST INFRASTRUCTURE ONLY.  Imported by tests/ as the checker of the GPU
network's forward and backward passes, never by the product path.

The network has no counterpart in the reference (it is the learned proposal
added to the hypothesis set of src/patchMatch.cu:827-830 / :935-977).  Its
operators are textbook: 3x3 convolution with zero padding 1 and stride 1 or 2
(NHWC activations, weights [Cout][3][3][Cin]), ReLU, nearest x2 upsampling
with concatenation, a sigmoid-parameterised inverse-depth head with L1 loss
and an error scale regressed (L1) onto the detached error.  bf16() emulates the GPU's rounding of
stored activations and gradients (round to nearest even) so that the two
passes can be compared step for step; pass round_bf16=False for the exact
FP64 network used to pin the backward pass by finite differences.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(64, 32, 1, 1), (32, 32, 1, 1), (32, 64, 2, 1), (64, 64, 1, 1), (64, 128, 2, 1),
          (128, 128, 1, 1), (128, 256, 2, 1), (256, 256, 1, 1), (384, 128, 1, 1),
          (128, 128, 1, 1), (192, 64, 1, 1), (64, 64, 1, 1), (96, 32, 1, 1), (32, 32, 1, 1),
          (32, 32, 1, 0)]
# input tensor of each layer: "x", a layer index, or a concatenation ("cat", up, skip)
INPUTS = ["x", 0, 1, 2, 3, 4, 5, 6, ("cat", 7, 5), 8, ("cat", 9, 3), 10, ("cat", 11, 1), 12, 13]


def bf16(a):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _cols(X, stride):
    N, H, W, C = X.shape
    Ho, Wo = (H - 1) // stride + 1, (W - 1) // stride + 1
    Xp = np.pad(X, ((0, 0), (1, 1), (1, 1), (0, 0)))
    cols = np.empty((N, Ho, Wo, 9, C))
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, 3 * ky + kx, :] = Xp[:, ky:ky + stride * (Ho - 1) + 1:stride,
                                               kx:kx + stride * (Wo - 1) + 1:stride, :]
    return cols


def conv_fwd(X, Wt, b, stride, relu):
    """Y[n,oy,ox,co] = act(b[co] + sum X[n, s oy + ky - 1, s ox + kx - 1, ci] Wt[co,ky,kx,ci])."""
    Cout = Wt.shape[0]
    cols = _cols(X, stride)
    Y = cols.reshape(-1, cols.shape[3] * cols.shape[4]) @ Wt.reshape(Cout, -1).T
    if b is not None:
        Y = Y + b
    Y = Y.reshape(cols.shape[:3] + (Cout,))
    return np.maximum(Y, 0.0) if relu else Y


def conv_bwd_data(dZ, Wt, H, W, stride):
    """Gradient of conv_fwd's pre-activation output dZ with respect to X (N x H x W x Cin)."""
    N, Ho, Wo, Cout = dZ.shape
    Cin = Wt.shape[3]
    dcols = (dZ.reshape(-1, Cout) @ Wt.reshape(Cout, -1)).reshape(N, Ho, Wo, 9, Cin)
    dXp = np.zeros((N, H + 2, W + 2, Cin))
    for ky in range(3):
        for kx in range(3):
            dXp[:, ky:ky + stride * (Ho - 1) + 1:stride,
                kx:kx + stride * (Wo - 1) + 1:stride, :] += dcols[:, :, :, 3 * ky + kx, :]
    return dXp[:, 1:H + 1, 1:W + 1, :]


def conv_bwd_weight(X, dZ, stride):
    """Gradient with respect to Wt: [Cout][3][3][Cin]."""
    Cout = dZ.shape[3]
    cols = _cols(X, stride)
    dW = dZ.reshape(-1, Cout).T @ cols.reshape(-1, cols.shape[3] * cols.shape[4])
    return dW.reshape(Cout, 3, 3, X.shape[3])


def upcat(A, S):
    """[nearest x2 of A cropped to S's size, S] along channels."""
    H, W = S.shape[1:3]
    up = A.repeat(2, axis=1).repeat(2, axis=2)[:, :H, :W, :]
    return np.concatenate([up, S], axis=3)


def upcat_bwd(dO, A_shape, Ca):
    N, Ha, Wa, _ = A_shape
    H, W = dO.shape[1:3]
    dA = np.zeros(A_shape)
    for dy in range(2):
        for dx in range(2):
            blk = dO[:, dy::2, dx::2, :Ca]
            dA[:, :blk.shape[1], :blk.shape[2], :] += blk
    return dA, dO[:, :, :, Ca:]


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


SCALE_LO, SCALE_HI = 1e-4, 0.5


def head_outputs(Hf, lo, hi):
    """Inverse depth u = lo + (hi - lo) sigmoid(z0) and the scale of its
    error b = SCALE_LO + (SCALE_HI - SCALE_LO) sigmoid(z1)."""
    u = lo + (hi - lo) * sigmoid(Hf[..., 0])
    b = SCALE_LO + (SCALE_HI - SCALE_LO) * sigmoid(Hf[..., 1])
    return u, b


def head_loss(Hf, tgt, lo, hi, w_conf, e_fixed=None):
    """L1 on u against tgt (> 0 only) plus w_conf x |b - |e|| with the error
    e = u - tgt held fixed (its gradient reaches b only), both summed over
    pixels / P.  e_fixed replaces e in the scale term (for finite differences
    of that surrogate).  Returns (dHf, mean L1, mean scale L1)."""
    P = Hf.shape[0] * Hf.shape[1] * Hf.shape[2]
    s0, s1 = sigmoid(Hf[..., 0]), sigmoid(Hf[..., 1])
    u, b = head_outputs(Hf, lo, hi)
    valid = tgt > 0
    e = u - tgt
    ae = np.abs(e if e_fixed is None else e_fixed)
    dH = np.zeros_like(Hf)
    dH[..., 0] = np.where(valid, np.sign(e) * (hi - lo) * s0 * (1 - s0) / P, 0.0)
    dH[..., 1] = np.where(valid, w_conf * np.sign(b - ae) * (SCALE_HI - SCALE_LO) * s1 * (1 - s1)
                          / P, 0.0)
    l1 = np.where(valid, np.abs(e), 0.0).sum() / P
    sl1 = np.where(valid, np.abs(b - ae), 0.0).sum() / P
    return dH, l1, sl1


def unpack(params, layers=LAYERS):
    """Split the flat parameter vector (all weights, then all biases)."""
    Ws, bs, o = [], [], 0
    for ci, co, _, _ in layers:
        Ws.append(np.asarray(params[o:o + co * 9 * ci], np.float64).reshape(co, 3, 3, ci))
        o += co * 9 * ci
    for _, co, _, _ in layers:
        bs.append(np.asarray(params[o:o + co], np.float64))
        o += co
    return Ws, bs


def forward(params, x, round_bf16=True):
    """Activations of every layer (dict) and the FP32-like head output."""
    rnd = bf16 if round_bf16 else (lambda a: a)
    Ws, bs = unpack(params)
    if round_bf16:
        Ws = [bf16(w) for w in Ws]
    act = {"x": np.asarray(x, np.float64)}
    for i, (ci, co, s, relu) in enumerate(LAYERS):
        src = INPUTS[i]
        if isinstance(src, tuple):
            key = f"cat{i}"
            act[key] = upcat(act[src[1]], act[src[2]])
            X = act[key]
        else:
            X = act[src]
        Y = conv_fwd(X, Ws[i], bs[i], s, relu)
        act[i] = Y if i == len(LAYERS) - 1 else rnd(Y)
    return act


def backward(params, act, tgt, lo, hi, w_conf, round_bf16=True):
    """Parameter gradients (flat, same layout as params) and the two losses."""
    rnd = bf16 if round_bf16 else (lambda a: a)
    Ws, _ = unpack(params)
    if round_bf16:
        Ws = [bf16(w) for w in Ws]
    nL = len(LAYERS)
    dH, l1, bce = head_loss(act[nL - 1], tgt, lo, hi, w_conf)
    grads_w, grads_b = [None] * nL, [None] * nL
    g = {nL - 1: rnd(dH)}
    for i in range(nL - 1, -1, -1):
        ci, co, s, relu = LAYERS[i]
        dZ = g[i]
        if relu:
            dZ = np.where(act[i] > 0, dZ, 0.0)
        src = INPUTS[i]
        X = act[f"cat{i}"] if isinstance(src, tuple) else act[src]
        grads_b[i] = dZ.reshape(-1, co).sum(axis=0)
        grads_w[i] = conv_bwd_weight(X, dZ, s)
        if src == "x":
            continue
        dX = conv_bwd_data(dZ, Ws[i], X.shape[1], X.shape[2], s)
        if isinstance(src, tuple):
            dX = rnd(dX)
            dA, dS = upcat_bwd(dX, act[src[1]].shape, LAYERS[src[1]][1])
            g[src[1]] = rnd(dA)
            g[src[2]] = dS
        elif src in g:
            # a skip tensor: its concatenation gradient came first
            g[src] = rnd(g[src] + dX)
        else:
            g[src] = rnd(dX)
    flat = np.concatenate([w.ravel() for w in grads_w] + [b for b in grads_b])
    return flat, l1, bce
