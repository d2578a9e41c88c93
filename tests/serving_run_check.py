"""A long serving run with its results checked (run on a GPU as a script,
not collected by pytest):

    python tests/serving_run_check.py --mode learned --rate 4 --duration 600

Replays the Poisson stream of patchmatch_cuda_amd.serving in one mode and
keeps the depth map of the first request of each resolution.  After the
stream, for each resolution:
* the kept depth map must equal a standalone run of the same request (FP16
  texels, one wave per sweep column, the same learned proposal) bit for bit;
* the same request's scene on a centred 384x288 crop (all its sources at
  full size), run the same way, is checked against the FP64 oracle running
  the same algorithm on the same FP16 images, generator states and
  proposal planes: every one of the run's sweeps teacher-forced (the oracle
  sweeps from the engine's state before it), decision for decision up to
  each column's first near-tie, with zero mismatches.  The end-to-end share
  of depths within 1% of the oracle's own run is reported beside it: a
  single near-tie decided the other way moves a whole propagation front, so
  that share is not a per-decision measure.
Prints one JSON line (serving figures and the checks) and exits non-zero if
a check fails.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from oracle import pmx_oracle as O  # noqa: E402
from patchmatch_cuda_amd import capi, configs, proposal, serving  # noqa: E402
from pmx_testlib import (centred_crop, compare_sweep, oracle_problem, proposal_planes,  # noqa: E402
                         to_frame, within)

CROP = (384, 288)


def crop_check(ctx, sc, n_src, rp, learned, weights, tau, prop_sweeps):
    """Teacher-forced and end-to-end oracle check of one request's scene on
    a centred crop; returns the check's record."""
    (x0, y0), (ref, srcs, cams, gt) = centred_crop(sc, *CROP)
    e = capi.Engine(ctx, rp, ref, list(srcs[:n_src]), cams[:n_src + 1])
    e.set_storage(capi.STORE_F16)
    e.set_sweep_waves(1)
    net = None
    try:
        e.prepare()
        d0, st0 = e.depth(), e.rng()
        prop = spread = None
        if learned:
            e.set_proposal_seeding(False)
            e.set_proposal_sweeps(prop_sweeps)
            W0, H0 = proposal.low_res(e.W, e.H)
            net = proposal.Net(ctx, 1, H0, W0, rp.depth_near, rp.depth_far)
            net.load(weights)
            prop, spread = proposal_planes(ctx, e, net, tau)
            e.propose(net, tau=tau)
        r16, s16 = e.images()
        pb = oracle_problem(r16, s16, e.homographies(), cams[0], rp)
        if learned:
            pb.set_proposal(prop.astype(np.float64), spread.astype(np.float64))
            pb.set_proposal_sweeps(prop_sweeps)
        t = time.perf_counter()
        e.cost_all_views(rotated=True)
        compared = total = mismatches = 0
        bad = []
        k = 0
        for it in range(rp.num_iterations):
            S = 1 if it == 0 else rp.num_samples   # the schedule of pmx_engine_run
            for down, rot in ((1, 1), (1, 0), (0, 1), (0, 0)):
                d_s = to_frame(e.depth().astype(np.float64), rot)
                c_s = to_frame(e.costs().astype(np.float64), rot)
                pb.set_sweep_id(k)
                ids_o, tr = pb.sweep(d_s, c_s, e.rng().copy(), down, rot, S, record_ids=True,
                                     trace=True)
                e.sweep(down, rot, S, record_ids=True)
                res = compare_sweep(to_frame(e.depth(), rot), d_s, to_frame(e.used_ids()[:S], rot),
                                    ids_o, tr, down)
                compared += res["compared"]
                total += res["total"]
                mismatches += res["mismatches"]
                if res["mismatches"]:
                    # the first mismatching decisions, in sweep-frame (row, column)
                    Hs = d_s.shape[0]
                    order = np.arange(1, Hs) if down else np.arange(Hs - 2, -1, -1)
                    dg_s, ig_s = to_frame(e.depth(), rot), to_frame(e.used_ids()[:S], rot)
                    for kk, c in res["first_bad"][:3]:
                        r = int(order[kk])
                        bad.append({"sweep": k, "down": down, "rotated": rot, "S": S, "row": r, "col": c,
                                    "depth_gpu": float(dg_s[r, c]), "depth_oracle": float(d_s[r, c]),
                                    "ids_gpu": ig_s[:, r, c].tolist(), "ids_oracle": ids_o[:, r, c].tolist(),
                                    "choice": int(tr["choice"][r, c]),
                                    "view_margin": float(tr["view_margin"][r, c]),
                                    "cost_margin": float(tr["cost_margin"][r, c])})
                k += 1
        dg = e.depth()
        pb.set_sweep_id(0)
        do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples, rp.num_iterations,
                          record_ids=False)
        return {"crop": f"{CROP[0]}x{CROP[1]} at ({x0}, {y0})", "sweeps": k,
                "teacher_forced_compared": compared, "teacher_forced_total": total,
                "teacher_forced_mismatches": mismatches, "first_mismatches": bad,
                "end_to_end_within_1pct_of_oracle": within(dg, do),
                "gt_within_1pct": {"gpu": within(dg, gt), "oracle": within(do.astype(np.float32), gt)},
                "oracle_s": time.perf_counter() - t}
    finally:
        if net is not None:
            net.close()
        e.close()


def log(*m):
    print(*m, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=("learned", "plain"), default="learned")
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--duration", type=float, default=600.0)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--tau", type=float, default=0.2)
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--oracle-threads", type=int, default=16)
    a = ap.parse_args()
    _, rp = configs.C2_REALISTIC
    t = time.perf_counter()
    pool = serving.scene_pool()
    log(f"scene pool rendered in {time.perf_counter() - t:.1f} s")
    stream = serving.make_stream(a.rate, a.duration, a.seed)
    first = {}
    for r in stream:
        first.setdefault(r.size, r)
    learned = a.mode == "learned"
    weights = proposal.DEFAULT_WEIGHTS if learned else None
    ctxs = [capi.Context(0), capi.Context(0)]
    srv = serving.Server(ctxs, rp, pool, a.budget_gib * 2 ** 30, weights, rp.num_iterations,
                         a.tau)
    srv.keep_results = {r.rid for r in first.values()}
    log(f"mode {a.mode}: {len(stream)} requests over {a.duration:.0f} s at {a.rate}/s")
    res = srv.run(stream, log)
    log(f"served: {json.dumps(res)}")
    ctx = ctxs[0]
    checks, ok = [], True
    O.set_threads(min(a.oracle_threads, os.cpu_count() or 1))
    for size in sorted(first):
        req = first[size]
        sc = pool[size]
        e = capi.Engine(ctx, rp, sc.ref, list(sc.srcs[:req.n_src]), sc.cams[:req.n_src + 1])
        e.set_storage(capi.STORE_F16)
        e.set_sweep_waves(1)
        net = None
        try:
            e.prepare()
            if learned:
                e.set_proposal_seeding(False)
                e.set_proposal_sweeps(srv.proposal_sweeps)
                W0, H0 = proposal.low_res(e.W, e.H)
                net = proposal.Net(ctx, 1, H0, W0, rp.depth_near, rp.depth_far)
                net.load(weights)
                e.propose(net, tau=a.tau)
            e.run()
            served = srv.results[req.rid]
            same = bool(np.array_equal(e.depth(), served))
            row = {"request": req.rid, "size": f"{e.W}x{e.H}", "sources": req.n_src,
                   "served_equals_standalone": same}
        finally:
            if net is not None:
                net.close()
            e.close()
        row["oracle_crop"] = crop_check(ctx, sc, req.n_src, rp, learned, weights, a.tau,
                                        srv.proposal_sweeps)
        oc = row["oracle_crop"]
        ok &= same and oc["teacher_forced_mismatches"] == 0 and \
            oc["teacher_forced_compared"] >= 0.25 * oc["teacher_forced_total"]
        checks.append(row)
        log(f"check: {row}")
    for c in ctxs:
        c.close()
    out = {"mode": a.mode, "rate_req_s": a.rate, "duration_s": a.duration, "seed": a.seed,
           "proposal_sweeps": srv.proposal_sweeps if learned else None,
           "requests": len(stream), "budget_gib": a.budget_gib, "served": res,
           "checks": checks, "checks_passed": bool(ok)}
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
