"""Learned depth proposal: the BF16 encoder-decoder of libpmx.so (pmx_net_*),
its training on the engine's own depth maps of seeded synthetic scenes, and
the weights file.

The network reads a 1/4-resolution plane-sweep cost volume (32 inverse-depth
planes on the two source views nearest the reference camera, made by
pmx_engine_cost_volume) plus the reference intensity and predicts inverse
depth and a confidence.  pmx_engine_propose
turns that into one extra depth hypothesis per pixel per sweep; the NCC cost
judges it with the other candidates (src/patchMatch.cu:613-618) and pixels
below the confidence threshold run plain PatchMatch.

The head predicts inverse depth u and the scale b of its error; the sweeps
draw one hypothesis per pixel per sweep from 1/depth +- 2b, and the
confidence 1 - exp(-0.01 u / b) gates where a proposal is made.

Training data: each scene is rendered (patchmatch_cuda_amd.scene), the plain
engine runs on it, and the 4x4-block mean inverse depth of its depth map is
the target: L1 on u plus 0.1 x L1 of b against the detached error |u - t|.
Adam, learning rate 1e-3 after a 500-step linear warm-up, batch 8, 100
epochs over 2000 scenes.

Device memory comes from the engine's allocator (capi.DeviceArray); every
call, copies included, is ordered on the context's stream.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import capi, configs

# (Cin, Cout, stride, relu) of the 15 convolutions, in execution order
LAYERS = [(64, 32, 1, 1), (32, 32, 1, 1), (32, 64, 2, 1), (64, 64, 1, 1), (64, 128, 2, 1),
          (128, 128, 1, 1), (128, 256, 2, 1), (256, 256, 1, 1), (384, 128, 1, 1),
          (128, 128, 1, 1), (192, 64, 1, 1), (64, 64, 1, 1), (96, 32, 1, 1), (32, 32, 1, 1),
          (32, 32, 1, 0)]
LAYER_NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "d2a", "d2b", "d1a",
               "d1b", "d0a", "d0b", "head"]
IN_C = 64
HEAD_C = 32
PLANES = 32

_vp = ctypes.c_void_p
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "proposal_bf16.pmxw")


def low_res(W, H):
    return (W + 3) // 4, (H + 3) // 4


class Net:
    """The proposal network on one context (pmx_net_*)."""
    CLOSE_ORDER = 1

    def __init__(self, ctx: capi.Context, max_batch: int, H0: int, W0: int, near: float,
                 far: float, seed: int = 0):
        self.ctx = ctx
        self.near, self.far = float(near), float(far)
        h = _vp()
        ctx.check(capi.lib().pmx_net_create(ctx.h, max_batch, H0, W0, near, far, seed,
                                            ctypes.byref(h)))
        self.h = h
        ctx.track(self)
        nt, nw = ctypes.c_size_t(), ctypes.c_size_t()
        ctx.check(capi.lib().pmx_net_num_params(h, ctypes.byref(nt), ctypes.byref(nw)))
        self.n_params, self.n_weights = nt.value, nw.value
        wo = np.zeros(len(LAYERS), np.uint64)
        bo = np.zeros(len(LAYERS), np.uint64)
        ctx.check(capi.lib().pmx_net_layer_offsets(h, _vp(wo.ctypes.data), _vp(bo.ctypes.data)))
        self.w_off, self.b_off = wo.astype(np.int64), bo.astype(np.int64)

    def close(self):
        if getattr(self, "h", None):
            capi.lib().pmx_net_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _c(self, rc):
        self.ctx.check(rc)

    # parameters
    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self._c(capi.lib().pmx_net_get_params(self.h, _vp(p.ctypes.data), p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._c(capi.lib().pmx_net_set_params(self.h, _vp(p.ctypes.data), p.size))

    def grads(self) -> np.ndarray:
        g = np.zeros(self.n_params, np.float32)
        self._c(capi.lib().pmx_net_get_grads(self.h, _vp(g.ctypes.data), g.size))
        return g

    def zero_grads(self):
        self._c(capi.lib().pmx_net_zero_grads(self.h))

    def layer_weight(self, p, i):
        ci, co, _, _ = LAYERS[i]
        o = int(self.w_off[i])
        return p[o:o + co * 9 * ci].reshape(co, 3, 3, ci)

    def layer_bias(self, p, i):
        o = int(self.b_off[i])
        return p[o:o + LAYERS[i][1]]

    # passes (x: DeviceArray N x H0 x W0 x 64 of bf16 bits; tgt: N x H0 x W0 float32)
    def forward(self, x):
        N, H0, W0, C = x.shape
        assert C == IN_C and x.dtype == np.uint16
        self._c(capi.lib().pmx_net_forward(self.h, _vp(x.ptr), N, H0, W0))

    def head(self, shape) -> np.ndarray:
        out = np.zeros(tuple(shape[:3]) + (HEAD_C,), np.float32)
        self._c(capi.lib().pmx_net_get_head(self.h, _vp(out.ctypes.data), out.size))
        return out

    def backward(self, x, tgt, w_conf=0.1):
        loss = np.zeros(2, np.float32)
        self._c(capi.lib().pmx_net_backward(self.h, _vp(x.ptr), _vp(tgt.ptr), w_conf,
                                            _vp(loss.ctypes.data)))
        return loss

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self._c(capi.lib().pmx_net_adam(self.h, lr, b1, b2, eps))

    def train_step(self, x, tgt, lr=1e-3, w_conf=0.1, want_loss=True):
        N, H0, W0, _ = x.shape
        loss = np.zeros(2, np.float32)
        self._c(capi.lib().pmx_net_train_step(self.h, _vp(x.ptr), _vp(tgt.ptr), N, H0, W0, lr,
                                              w_conf,
                                              _vp(loss.ctypes.data) if want_loss else None))
        return loss

    def propose(self, W, H, tau, depth, spread=None, conf=None):
        """depth, spread, conf: DeviceArray H x W float32."""
        opt = lambda a: (_vp(a.ptr), W * 4) if a is not None else (None, 0)  # noqa: E731
        self._c(capi.lib().pmx_net_propose(self.h, W, H, tau, _vp(depth.ptr), W * 4, *opt(spread),
                                           *opt(conf)))

    # weights file: bf16 weights + FP32 biases, float32 .npy (no pickle)
    def save(self, path):
        p = self.params()
        w = p[:self.n_weights].view(np.uint32)
        wb = ((w + 0x7FFF + ((w >> 16) & 1)) >> 16).astype(np.uint16)
        meta = np.frombuffer(json.dumps({"layers": LAYERS, "near": self.near,
                                         "far": self.far}).encode(), np.uint8)
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "wb") as f:
            np.save(f, wb, allow_pickle=False)
            np.save(f, p[self.n_weights:].astype(np.float32), allow_pickle=False)
            np.save(f, meta, allow_pickle=False)

    def load(self, path):
        with open(path, "rb") as f:
            wb = np.load(f, allow_pickle=False)
            b = np.load(f, allow_pickle=False)
            meta = json.loads(bytes(np.load(f, allow_pickle=False)).decode())
        if [tuple(x) for x in meta["layers"]] != LAYERS or wb.size != self.n_weights:
            raise ValueError(f"{path}: weights of a different network")
        w = (wb.astype(np.uint32) << 16).view(np.float32)
        self.set_params(np.concatenate([w, b]))


# ------------------------------------------------------------ training data
def training_spec(i, W=640, H=480, V=6, seed0=10_000, focal=None):
    """Scene i of the training set: the 2048x1536 workload's world statistics
    at 640x480, 6 sources, its own seeds.  focal None scales the focal length
    with the width (the same field of view); a fixed focal length keeps the
    2048x1536 view's pixel disparities in a narrower field of view."""
    f = 1800.0 * W / 2048.0 if focal is None else float(focal)
    return configs.SceneSpec(W, H, V, 20, 5, 1.0, 20.0, f, 0.1, 1.0, 15.0,
                             0.01, seed0 + i, seed0 + 50_000 + i)


def training_params(V=6):
    _, rp = configs.C2_REALISTIC
    return rp.replace(num_samples=(V + 1) // 2)


def make_dataset(ctx, n, W=640, H=480, V=6, seed0=10_000, threads=16, log=print,
                 views=None, focal=None):
    """Renders n scenes, runs the plain engine on each and returns device
    arrays x (n, H0, W0, 64) bf16 and tgt (n, H0, W0) float32."""
    from . import scene as scene_mod
    W0, H0 = low_res(W, H)
    x_all = capi.DeviceArray(ctx, (n, H0, W0, IN_C), np.uint16)
    t_all = capi.DeviceArray(ctx, (n, H0, W0), np.float32)
    depth = capi.DeviceArray(ctx, (H, W), np.float32)
    rp = training_params(V)
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        futs = [pool.submit(scene_mod.render, training_spec(i, W, H, V, seed0, focal))
                for i in range(n)]
        for i, fut in enumerate(futs):
            sc = fut.result()
            futs[i] = None
            eng = capi.Engine(ctx, rp, sc.ref, sc.srcs, sc.cams)
            eng.prepare()
            eng.run()
            eng.cost_volume(*(views(sc) if views else eng.volume_views()), x_all.at(i))
            depth.upload(eng.depth())
            ctx.check(capi.lib().pmx_cnn_target(ctx.h, _vp(depth.ptr), W * 4, W, H,
                                                _vp(t_all.at(i))))
            ctx.sync()
            eng.close()
            if (i + 1) % 100 == 0 or i + 1 == n:
                log(f"dataset {i + 1}/{n} scenes, {time.time() - t0:.1f} s")
    return x_all, t_all


def train(net, x_all, t_all, epochs, batch=8, lr=1e-3, w_conf=0.1, seed=0, log=print,
          warmup_steps=500):
    """Adam at learning rate lr after a linear warm-up over warmup_steps
    steps.  Without it the first, sign-like Adam steps move every weight of
    this normalisation-free network by lr at once, which scales the
    activations of 15 stacked convolutions by orders of magnitude and
    saturates the head's sigmoids (gradients then vanish for good)."""
    n = x_all.shape[0]
    rng = np.random.default_rng(seed)
    xb = capi.DeviceArray(net.ctx, (batch,) + tuple(x_all.shape[1:]), x_all.dtype)
    tb = capi.DeviceArray(net.ctx, (batch,) + tuple(t_all.shape[1:]), t_all.dtype)
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        order = rng.permutation(n)
        sums = np.zeros(2)
        steps = 0
        for s in range(0, n - batch + 1, batch):
            for j, k in enumerate(order[s:s + batch]):
                xb.copy_rows(j, x_all, int(k))
                tb.copy_rows(j, t_all, int(k))
            it = ep * (n // batch) + steps
            lr_t = lr * min(1.0, (it + 1) / warmup_steps) if warmup_steps > 0 else lr
            loss = net.train_step(xb, tb, lr_t, w_conf)
            sums += loss
            steps += 1
        hist.append((sums / max(steps, 1)).tolist())
        log(f"epoch {ep + 1}/{epochs}: L1 on inverse depth {hist[-1][0]:.5f}, "
            f"L1 of the error scale {hist[-1][1]:.5f} ({time.time() - t0:.1f} s)")
    return hist


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the proposal network on engine depth maps")
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--warmup-steps", type=int, default=500)
    a = ap.parse_args(argv)
    ctx = capi.Context(0)
    W, H, V = 640, 480, 6
    W0, H0 = low_res(W, H)
    log = lambda *m: print(*m, flush=True)  # noqa: E731
    x_all, t_all = make_dataset(ctx, a.scenes, W, H, V, log=log)
    _, rp = configs.C2_REALISTIC
    net = Net(ctx, a.batch, H0, W0, rp.depth_near, rp.depth_far, seed=a.seed)
    hist = train(net, x_all, t_all, a.epochs, a.batch, a.lr, seed=a.seed, log=log,
                 warmup_steps=a.warmup_steps)
    net.save(a.out)
    log(json.dumps({"weights": a.out, "scenes": a.scenes, "epochs": a.epochs, "loss": hist}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
