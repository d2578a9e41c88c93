"""Serving loop: a seeded Poisson stream of MVS requests run on one GPU.

Each request is one reference view with 4-20 source views at one of five
resolutions from 640x480 to 3072x2048, FP16 texels, and (in the learned
mode) the learned proposal.  The loop admits waiting requests in arrival
order while the device bytes of the resident views (pmx_engine_footprint)
stay within an HBM budget, keeps each admitted view's images and state
resident on the device from upload to result, and runs everything admitted
as one batch (pmx_batch_run: one persistent sweep launch per sweep step
over all views).  While a batch runs, the next group of arrivals is
uploaded, prepared and proposed on a second context (its own HIP stream).
Latency is arrival -> depth map on the host.

The reference serves nothing (one reference view per process,
src/main.cpp:96-127); this is the request path around its runPatchMatch
(src/patchMatch.cu:263-407).

Scenes are rendered before the loop starts: one 20-source scene per
resolution, of which a request with V sources uses the first V (the
rendering cost would otherwise be part of the measured latency).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import capi, configs, proposal, scene

SIZES = [(640, 480), (1024, 768), (1600, 1200), (2048, 1536), (3072, 2048)]
MAX_SRC = 20


@dataclass
class Request:
    rid: int
    arrival: float       # seconds from the start of the loop
    size: int            # index into SIZES
    n_src: int


def make_stream(rate, duration, seed=11, min_src=4, max_src=MAX_SRC):
    """Poisson arrivals at `rate` per second over `duration` seconds."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out
        out.append(Request(len(out), t, int(rng.integers(len(SIZES))),
                           int(rng.integers(min_src, max_src + 1))))


def scene_pool(seed=21):
    """One rendered 20-source scene per resolution, with the world statistics
    of the 2048x1536 workload."""
    pool = []
    for i, (W, H) in enumerate(SIZES):
        spec = configs.SceneSpec(W, H, MAX_SRC, 20, 5, 1.0, 20.0, 1800.0 * W / 2048.0, 0.1, 1.0,
                                 15.0, 0.01, seed + i, seed + 100 + i)
        pool.append(scene.render(spec))
    return pool


def _pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else float("nan")


class Server:
    """Runs a request stream.  Two contexts (two HIP streams) alternate: while
    the batch admitted last runs on one, the next group of waiting requests
    is uploaded, prepared and proposed on the other, so host-side admission
    work overlaps the sweeps."""

    def __init__(self, ctxs, rp, pool, budget_bytes, learned_weights=None,
                 learned_iterations=None, tau=0.2, seeding=False, max_batch=32,
                 proposal_sweeps=4):
        self.ctxs, self.rp, self.pool = list(ctxs), rp, pool
        self.budget = int(budget_bytes)
        self.weights = learned_weights
        self.iters = learned_iterations or rp.num_iterations
        self.tau = tau
        self.seeding = seeding
        # the proposal joins the first iteration's sweeps only (as accurate
        # as every sweep on the 2048x1536 workload, at a quarter of the cost)
        self.proposal_sweeps = proposal_sweeps
        self.max_batch = max_batch
        self.nets = {}
        self.resident = 0
        self.keep_results = False   # True: every depth map; a set: those request ids
        self.results = {}
        self.launch_shapes = []   # batch launch shape of every group of the last run

    def dims(self, req):
        spec = self.pool[req.size].spec
        return spec.W, spec.H

    def footprint(self, req):
        W, H = self.dims(req)
        return capi.engine_footprint(self.rp, W, H, [(W, H)] * req.n_src, capi.STORE_F16,
                                     proposal=self.weights is not None)

    def _net_for(self, ctx, W, H):
        """The proposal network for one context and resolution."""
        key = (id(ctx), W, H)
        if key not in self.nets:
            W0, H0 = proposal.low_res(W, H)
            n = proposal.Net(ctx, 1, H0, W0, self.rp.depth_near, self.rp.depth_far)
            n.load(self.weights)
            self.nets[key] = n
        return self.nets[key]

    def _stage(self, ctx, waiting):
        """Admits waiting requests in arrival order within the HBM budget
        (at least one), uploads them on ctx, prepares them and installs the
        learned proposal.  Returns the staged group or None."""
        if not waiting:
            return None
        admitted = []
        while waiting and len(admitted) < self.max_batch:
            fp = self.footprint(waiting[0])
            if self.resident + fp > self.budget and (admitted or self.resident > 0):
                break
            req = waiting.popleft()
            self.resident += fp
            sc = self.pool[req.size]
            eng = capi.Engine(ctx, self.rp, sc.ref, list(sc.srcs[:req.n_src]),
                              sc.cams[:req.n_src + 1])
            eng.set_storage(capi.STORE_F16)
            if self.weights is not None:
                eng.set_num_iterations(self.iters)
                eng.set_proposal_seeding(self.seeding)
                eng.set_proposal_sweeps(self.proposal_sweeps)
            admitted.append((req, fp, eng))
        if not admitted:
            return None
        b = capi.Batch(ctx, [e for _, _, e in admitted])
        b.set_graph(False)
        # launch shape 3 (a third resident wave per SIMD) pays once the group
        # has more column strips than shape 1 keeps resident; with fewer its
        # shorter tap batches only add latency
        strips = sum(min(self.dims(req)) for req, _, _ in admitted)
        if strips > b.launch()[1]:
            b.set_launch(3, 0)
        self.launch_shapes.append(b.launch()[0])
        b.prepare()
        if self.weights is not None:
            for req, _, e in admitted:
                e.propose(self._net_for(ctx, *self.dims(req)), tau=self.tau)
        return {"batch": b, "admitted": admitted}

    def _finish(self, st, t0, lat):
        """Waits for a launched group, reads every depth map back, frees it."""
        px = 0
        for req, fp, e in st["admitted"]:
            d = e.depth()                # synchronises the group's stream
            keep = self.keep_results
            if keep is True or (keep and req.rid in keep):
                self.results[req.rid] = d
            lat.append(time.perf_counter() - t0 - req.arrival)
            W, H = self.dims(req)
            px += W * H
        st["batch"].close()
        for req, fp, e in st["admitted"]:
            e.close()
            self.resident -= fp
        return px

    def run(self, stream, log=print):
        self.launch_shapes = []
        pending = deque(stream)
        waiting = deque()
        lat, done_px, sizes = [], 0, []
        t0 = time.perf_counter()
        last_log = 0.0
        n_total = len(stream)
        turn = 0
        staged = None

        def arrive():
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            return now

        while pending or waiting or staged:
            now = arrive()
            if staged is None:
                if not waiting:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.05)))
                    continue
                staged = self._stage(self.ctxs[turn % len(self.ctxs)], waiting)
                turn += 1
            running = staged
            running["batch"].run()           # enqueued; returns at once
            arrive()
            staged = self._stage(self.ctxs[turn % len(self.ctxs)], waiting)
            if staged is not None:
                turn += 1
            done_px += self._finish(running, t0, lat)
            sizes.append(len(running["admitted"]))
            now = time.perf_counter() - t0
            if now - last_log > 10.0:
                log(f"t={now:.0f}s: {len(lat)}/{n_total} done, {len(waiting)} waiting, "
                    f"p50 latency {_pct(lat, 50):.2f}s")
                last_log = now
        wall = time.perf_counter() - t0
        for n in self.nets.values():
            n.close()
        self.nets.clear()
        return {
            "requests": len(lat),
            "wall_s": wall,
            "throughput_req_s": len(lat) / wall,
            "throughput_mpix_s": done_px / wall / 1e6,
            "latency_s": {"p50": _pct(lat, 50), "p95": _pct(lat, 95), "p99": _pct(lat, 99),
                          "max": float(max(lat)) if lat else float("nan")},
            "batches": len(sizes),
            "mean_batch": float(np.mean(sizes)) if sizes else 0.0,
            "groups_by_launch_shape": {str(k): self.launch_shapes.count(k)
                                       for k in sorted(set(self.launch_shapes))},
        }


def main(argv=None):
    ap = argparse.ArgumentParser(description="Poisson-stream serving loop (learned vs plain)")
    ap.add_argument("--rate", type=float, default=4.0, help="requests per second")
    ap.add_argument("--duration", type=float, default=600.0, help="seconds of arrivals")
    ap.add_argument("--budget-gib", type=float, default=64.0, help="HBM budget for views")
    ap.add_argument("--learned-iterations", type=int, default=4)
    ap.add_argument("--tau", type=float, default=0.2)
    ap.add_argument("--proposal-sweeps", type=int, default=4,
                    help="sweeps of a run that get the proposal (0: all)")
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--weights", default=proposal.DEFAULT_WEIGHTS)
    ap.add_argument("--modes", default="learned,plain")
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)  # noqa: E731
    _, rp = configs.C2_REALISTIC
    t = time.perf_counter()
    pool = scene_pool()
    log(f"scene pool rendered in {time.perf_counter() - t:.1f} s")
    stream = make_stream(a.rate, a.duration, a.seed)
    ctxs = [capi.Context(0), capi.Context(0)]
    out = {"rate_req_s": a.rate, "duration_s": a.duration, "budget_gib": a.budget_gib,
           "requests": len(stream), "seed": a.seed,
           "sizes": [f"{w}x{h}" for w, h in SIZES], "sources": "4-20"}
    for mode in a.modes.split(","):
        srv = Server(ctxs, rp, pool, a.budget_gib * 2 ** 30,
                     a.weights if mode == "learned" else None, a.learned_iterations, a.tau,
                     proposal_sweeps=a.proposal_sweeps)
        log(f"mode {mode}: {len(stream)} requests over {a.duration:.0f} s")
        res = srv.run(stream, log)
        res["iterations"] = srv.iters
        if mode == "learned":
            res["proposal_sweeps"] = srv.proposal_sweeps
        out[mode] = res
        log(f"mode {mode}: {json.dumps(res)}")
    for c in ctxs:
        c.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
