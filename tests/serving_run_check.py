"""A long serving run with its results checked (run on a GPU as a script,
not collected by pytest):

    python tests/serving_run_check.py --mode learned --rate 4 --duration 600

Replays the Poisson stream of patchmatch_cuda_amd.serving in one mode and
keeps the depth map of the first request of each resolution.  After the
stream: each kept depth map must equal a standalone run of the same request
(FP16 texels, one wave per sweep column, the same learned proposal) bit for
bit, and the 640x480 one must have >= 99% of its depths within 1% of the
FP64 oracle running the same algorithm on the same FP16 images, generator
states and proposal planes.  Prints one JSON line (serving figures and the
checks) and exits non-zero if a check fails.
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
from pmx_testlib import oracle_problem, proposal_planes, within  # noqa: E402


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
            d0, st0 = e.depth(), e.rng()
            prop = spread = None
            if learned:
                e.set_proposal_seeding(False)
                e.set_proposal_sweeps(srv.proposal_sweeps)
                W0, H0 = proposal.low_res(e.W, e.H)
                net = proposal.Net(ctx, 1, H0, W0, rp.depth_near, rp.depth_far)
                net.load(weights)
                if size == 0:
                    prop, spread = proposal_planes(ctx, e, net, a.tau)
                e.propose(net, tau=a.tau)
            e.run()
            served = srv.results[req.rid]
            same = bool(np.array_equal(e.depth(), served))
            row = {"request": req.rid, "size": f"{e.W}x{e.H}", "sources": req.n_src,
                   "served_equals_standalone": same}
            ok &= same
            if size == 0:
                r16, s16 = e.images()
                pb = oracle_problem(r16, s16, e.homographies(), sc.cams[0], rp)
                if learned:
                    pb.set_proposal(prop.astype(np.float64), spread.astype(np.float64))
                    pb.set_proposal_sweeps(srv.proposal_sweeps)
                t = time.perf_counter()
                do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples,
                                  rp.num_iterations, record_ids=False)
                agree = within(served, do)
                row["within_1pct_of_oracle"] = agree
                row["oracle_s"] = time.perf_counter() - t
                ok &= agree >= 0.99
            checks.append(row)
            log(f"check: {row}")
        finally:
            if net is not None:
                net.close()
            e.close()
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
