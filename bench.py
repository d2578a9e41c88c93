#!/usr/bin/env python3
"""Depth throughput of the MI355X PatchMatch engine.

One step = one full PatchMatch run of one reference view against its source
views with the images already resident in HBM: pre-blur, reference box
sums, RNG init, initial depth, cost of all views, and every iteration's
four propagation sweeps with the transposes between them
(PatchMatch::PatchMatch + runPatchMatch of the reference).

Workload (default): 1 reference + 10 source views at 2048x1536, float32,
window 15x15, 4 iterations, S = 5, seeded synthetic scene of 20 planes and
5 spheres.  value = reference pixels x steps / wall time, in Mpix/s.

Also printed (same JSON line):
  roofline      the sweep kernel's achieved FP32 rate: 21 algorithmic FLOPs
                per window sample x samples counted on the device / the
                average sweep launch time from HIP events on the engine's
                stream, against MI355X's 157.3 TFLOP/s FP32 peak; traffic =
                HBM bytes per launch from the committed rocprofv3 counter
                summary of the same workload (profiles/).
  cpu_baseline  the FP64 oracle (OpenMP) on a crop of the same workload.
  accuracy      the GPU engine on the same crop against the oracle: share of
                final depths within 1% relative.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("depth Mpix/s (10 src views) at ≥99% px within 1% of FP64 oracle; "
          "learned vs plain")
FLOPS_PER_SAMPLE = 21.0
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_sample(s):
    w, h = s.lower().split("x")
    return int(w), int(h)


def profiled_traffic(workload):
    """HBM bytes per sweep launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE,
    scaled by a copy kernel) and recorded in profiles/; None if not profiled."""
    path = os.path.join(ROOT, "profiles", f"sweep_kernel_{workload}.md")
    try:
        with open(path) as f:
            for line in f:
                if "traffic_bytes_per_launch:" in line:
                    tok = line.split("traffic_bytes_per_launch:")[1].strip().split(")")[0]
                    return float(tok), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def run_gpu_steps(eng, steps, warmup):
    from patchmatch_cuda_amd import capi  # noqa: F401

    for i in range(warmup):
        eng.prepare()
        eng.run()
        log(f"warmup {i + 1}/{warmup} done")
    eng.ctx.sync()
    eng.reset_stats()
    eng.set_timing(True)
    t0 = time.perf_counter()
    for i in range(steps):
        eng.prepare()
        eng.run()
    eng.ctx.sync()
    t1 = time.perf_counter()
    eng.set_timing(False)
    stats = eng.stats()
    log(f"timed {steps} steps: {(t1 - t0) * 1e3 / steps:.2f} ms/step")
    return (t1 - t0), stats


def proposal_planes(ctx, eng, net, tau):
    """The network's proposal for the engine's (prepared) view, as host
    planes: depth (0 = none) and spread in inverse depth, W x H."""
    from patchmatch_cuda_amd import capi, proposal

    W, H = eng.W, eng.H
    W0, H0 = proposal.low_res(W, H)
    x = capi.DeviceArray(ctx, (1, H0, W0, proposal.IN_C), np.uint16)
    d = capi.DeviceArray(ctx, (H, W), np.float32)
    sp = capi.DeviceArray(ctx, (H, W), np.float32)
    eng.cost_volume(*eng.volume_views(), x.ptr)
    net.forward(x)
    net.propose(W, H, tau, d, sp)
    out = d.download(), sp.download()
    for a in (x, d, sp):
        a.close()
    return out


def crop_problem(sc, sample):
    """The centred crop of the reference view (all sources kept)."""
    from patchmatch_cuda_amd import scene

    W, H = sc.spec.W, sc.spec.H
    w, h = sample
    x0, y0 = (W - w) // 2, (H - h) // 2
    return (x0, y0, w, h), scene.crop_reference(sc, x0, y0, w, h)


def learned_leg(ctx, eng, sc, rp, weights, d_oracle, sample, parity_crop, steps, warmup, tau,
                seeding, prop_sweeps):
    """Learned-proposal PatchMatch against plain PatchMatch.

    Sweeps needed: on the view the accuracy oracle covered (sample None: the
    whole view; else its centred crop), each method runs 1..iterations
    iterations, and its final depth is compared with the converged FP64
    oracle (plain method, all iterations, d_oracle) and with the ground
    truth.  A method's count is the smallest whose depth has >= 99% of
    pixels within 1% of the oracle; the ground-truth count is the smallest
    that comes within 0.1 points of the plain method's converged share of
    pixels within 1% of the ground truth (a method that converges to a
    different but equally good solution meets the second, not the first).
    The learned method is timed on the full view at its count (a step =
    prepare + cost volume + network + proposal + run), and its parity is
    checked on the centred parity_crop: the GPU's learned run against the
    FP64 oracle running the same learned algorithm (same proposal planes,
    generator states and initial depth)."""
    from oracle import pmx_oracle as O
    from patchmatch_cuda_amd import capi, proposal

    W0, H0 = proposal.low_res(eng.W, eng.H)
    net = proposal.Net(ctx, 1, H0, W0, rp.depth_near, rp.depth_far)
    net.load(weights)
    eng.set_proposal_seeding(seeding)
    eng.set_proposal_sweeps(prop_sweeps)
    eng.prepare()
    prop, spread = proposal_planes(ctx, eng, net, tau)
    eng.set_proposal(None, 0)
    if sample is None:
        seng, gt, spd, ssd = eng, sc.gt_depth, None, None
        where = f"whole {eng.W}x{eng.H} view"
    else:
        (x0, y0, w, h), (ref, srcs, cams, gt) = crop_problem(sc, sample)
        seng = capi.Engine(ctx, rp, ref, list(srcs), cams)
        seng.set_proposal_seeding(seeding)
        seng.set_proposal_sweeps(prop_sweeps)
        spd = capi.DeviceArray(ctx, (h, w), np.float32).upload(
            np.ascontiguousarray(prop[y0:y0 + h, x0:x0 + w]))
        ssd = capi.DeviceArray(ctx, (h, w), np.float32).upload(
            np.ascontiguousarray(spread[y0:y0 + h, x0:x0 + w]))
        where = f"centred {w}x{h} crop"
    agree = lambda d: float(np.mean(np.abs(d - d_oracle) <= 0.01 * np.abs(d_oracle)))  # noqa
    on_gt = lambda d: float(np.mean(np.abs(d - gt) <= 0.01 * gt))  # noqa: E731
    study = {"plain": {}, "learned": {}}
    for k in range(1, rp.num_iterations + 1):
        seng.set_num_iterations(k)
        for method in ("plain", "learned"):
            seng.prepare()
            if method == "learned":
                if spd is None:
                    seng.propose(net, tau=tau)
                else:
                    seng.set_proposal(spd.ptr, spd.shape[1] * 4, ssd.ptr)
            seng.run()
            d = seng.depth().astype(np.float64)
            seng.set_proposal(None, 0)
            study[method][k] = {"sweeps": 4 * k,
                                "within_1pct_of_converged_oracle": agree(d),
                                "within_1pct_of_ground_truth": on_gt(d)}
        log(f"{where}, {k} iteration(s): plain {study['plain'][k]['within_1pct_of_converged_oracle']:.4f}"
            f" / {study['plain'][k]['within_1pct_of_ground_truth']:.4f}, learned "
            f"{study['learned'][k]['within_1pct_of_converged_oracle']:.4f} / "
            f"{study['learned'][k]['within_1pct_of_ground_truth']:.4f} of depths within 1% of the "
            f"converged oracle / of the ground truth")
    seng.set_num_iterations(rp.num_iterations)
    if spd is not None:
        seng.close()
        spd.close()
        ssd.close()
    gt_final = study["plain"][rp.num_iterations]["within_1pct_of_ground_truth"]
    need, need_gt = {}, {}
    for m in ("plain", "learned"):
        need[m] = next((k for k in sorted(study[m])
                        if study[m][k]["within_1pct_of_converged_oracle"] >= 0.99), None)
        need_gt[m] = next((k for k in sorted(study[m])
                           if study[m][k]["within_1pct_of_ground_truth"] >= gt_final - 0.001), None)
    k = need["learned"] or rp.num_iterations
    # parity of the learned path at k, on the parity crop
    (x0, y0, w, h), (ref, srcs, cams, gtc) = crop_problem(sc, parity_crop)
    pc = np.ascontiguousarray(prop[y0:y0 + h, x0:x0 + w])
    sc_ = np.ascontiguousarray(spread[y0:y0 + h, x0:x0 + w])
    ceng = capi.Engine(ctx, rp, ref, list(srcs), cams)
    ceng.set_proposal_seeding(seeding)
    ceng.set_proposal_sweeps(prop_sweeps)
    pd = capi.DeviceArray(ctx, (h, w), np.float32).upload(pc)
    sd = capi.DeviceArray(ctx, (h, w), np.float32).upload(sc_)
    ceng.set_num_iterations(k)
    ceng.prepare()
    ceng.set_proposal(pd.ptr, w * 4, sd.ptr)
    d0, st0 = ceng.depth(), ceng.rng()
    ceng.run()
    dg = ceng.depth().astype(np.float64)
    ref_b, srcs_b = ceng.images()
    f = lambda x: float(np.float32(x))  # noqa: E731
    pb = O.Problem(ref_b, srcs_b, ceng.homographies(), rp.half_window, f(rp.depth_near),
                   f(rp.depth_far), f(rp.spm_alpha), O.inv_k4(cams[0, :9]),
                   f(rp.orientation_x), f(rp.orientation_z))
    pb.set_proposal(pc.astype(np.float64), sc_.astype(np.float64))
    pb.set_proposal_sweeps(prop_sweeps)
    do, _, _ = pb.run(d0.astype(np.float64), st0.copy(), rp.num_samples, k, record_ids=False)
    parity = float(np.mean(np.abs(dg - do) <= 0.01 * np.abs(do)))
    log(f"learned on the {w}x{h} crop, {k} iteration(s): {parity:.4f} of depths within 1% of the "
        f"oracle running the same learned algorithm")
    ceng.close()
    pd.close()
    sd.close()

    # full-view timing at k (learned) and at the plain method's own count
    def timed(iters, learned):
        eng.set_num_iterations(iters)
        for i in range(warmup + steps):
            if i == warmup:
                ctx.sync()
                t0 = time.perf_counter()
            eng.prepare()
            if learned:
                eng.propose(net, tau=tau)
            eng.run()
        ctx.sync()
        wall = time.perf_counter() - t0
        gt_share = float(np.mean(np.abs(eng.depth() - sc.gt_depth) <= 0.01 * sc.gt_depth))
        eng.set_proposal(None, 0)
        eng.set_num_iterations(rp.num_iterations)
        return wall, gt_share

    wall, full_gt = timed(k, True)
    kp = need["plain"] or rp.num_iterations
    wall_p, _ = timed(kp, False)
    net.close()
    log(f"learned: {k} iteration(s), {wall * 1e3 / steps:.2f} ms/step; plain at {kp} "
        f"iteration(s): {wall_p * 1e3 / steps:.2f} ms/step")
    return {
        "iterations": k,
        "sweeps": 4 * k,
        "ms_per_step": wall * 1e3 / steps,
        "value": eng.W * eng.H * steps / wall / 1e6,
        "proposal_coverage": float(np.mean(prop > 0)),
        "plain_at_its_99pct_count": {"iterations": kp,
                                     "ms_per_step": wall_p * 1e3 / steps,
                                     "value": eng.W * eng.H * steps / wall_p / 1e6},
        "speedup_vs_plain_at_equal_accuracy": wall_p / wall,
        "within_1pct_of_ground_truth": full_gt,
        "iterations_to_99pct_of_converged_oracle": need,
        "iterations_to_plain_converged_ground_truth_accuracy": need_gt,
        "sweep_study": study,
        "sweep_study_sample": where,
        "parity": {"within_1pct_of_fp64_oracle": parity,
                   "sample": f"{w}x{h} centred crop, {k} iteration(s), oracle given the same "
                             f"proposal planes"},
        "tau": tau,
        "proposal_seeds_depth": bool(seeding),
        "proposal_sweeps": prop_sweeps,
        "weights": os.path.relpath(weights, ROOT),
    }


def batch_bench(args):
    """The batched workload: 16 reference views of mixed resolution, FP16
    texels, one persistent sweep launch per sweep step over all views,
    replayed from a hipGraph, launch shape from the autotuner.  Compared
    with the same views run one at a time (FP32 and FP16).  Per-view error:
    each view's centred crop, run in a batch of crops, against the FP64
    oracle on the same FP16-rounded images (parity) and on the FP32 images
    (the precision cost)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import pmx_oracle as O
    from patchmatch_cuda_amd import capi, configs, scene

    views = configs.c3_batch()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=4) as pool:
        scenes = list(pool.map(lambda v: scene.render(v[0]), views))
    log(f"rendered {len(views)} views in {time.perf_counter() - t0:.1f} s")
    ctx = capi.Context(0)
    engs = [capi.Engine(ctx, rp, sc.ref, list(sc.srcs), sc.cams) for (_, rp), sc in
            zip(views, scenes)]
    npx = sum(e.W * e.H for e in engs)

    def one_at_a_time(storage):
        for e in engs:
            e.set_storage(storage)
        for rep in range(2):             # warm-up pass, then the timed pass
            ctx.sync()
            t = time.perf_counter()
            for e in engs:
                e.prepare()
                e.run()
            ctx.sync()
            dt = time.perf_counter() - t
        return npx / dt / 1e6, dt * 1e3

    seq32 = seq32_ms = seq16 = seq16_ms = None
    if not args.no_cpu_baseline:
        seq32, seq32_ms = one_at_a_time(capi.STORE_F32)
        log(f"one view at a time, FP32: {seq32:.2f} Mpix/s")
        seq16, seq16_ms = one_at_a_time(capi.STORE_F16)
        log(f"one view at a time, FP16: {seq16:.2f} Mpix/s")
    else:
        for e in engs:
            e.set_storage(capi.STORE_F16)
    b = capi.Batch(ctx, engs)
    if args.batch_shape is not None:
        b.set_launch(args.batch_shape, 0)
        waves, wgs = b.launch()
    elif args.no_autotune:
        waves, wgs = b.launch()
    else:
        waves, wgs, tune_ms = b.autotune()
        log(f"autotuned: launch shape {waves}, {wgs} workgroups ({tune_ms:.1f} ms)")
    b.set_graph(True)
    for _ in range(args.warmup):
        b.prepare()
        b.run()
    ctx.sync()
    run_ms = 0.0
    t = time.perf_counter()
    for _ in range(args.steps):
        b.prepare()
        run_ms += b.run(timed=True)
    ctx.sync()
    wall = time.perf_counter() - t
    value = npx * args.steps / wall / 1e6
    log(f"batched: {value:.2f} Mpix/s ({wall * 1e3 / args.steps:.1f} ms/step)")
    b.close()
    for e in engs:
        e.close()
    line = {
        "metric": "batched depth Mpix/s (16 views, 10 src views each, mixed resolution)",
        "value": value,
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 texels, f32 sums, f64 homographies",
        "data": "synthetic (16 seeded scenes, no dataset)",
        "config": {"workload": "c3_batch", "views": [f"{e[0].W}x{e[0].H}" for e in views],
                   "src_views": 10, "window": 15, "iterations": 4, "num_samples": 5,
                   "graph": True, "launch_shape": waves, "workgroups": wgs},
        "gpu_ms_per_run": run_ms / args.steps,
    }
    if args.no_cpu_baseline:
        ctx.close()
        return line
    line["one_view_at_a_time"] = {"fp32_mpix_s": seq32, "fp32_ms": seq32_ms,
                                  "fp16_mpix_s": seq16, "fp16_ms": seq16_ms}
    line["speedup_vs_one_at_a_time_fp32"] = value / seq32

    # per-view error on centred crops
    cw, ch = parse_sample(args.batch_crop)
    crops, cengs = [], []
    for (spec, rp), sc in zip(views, scenes):
        x0, y0 = (spec.W - cw) // 2, (spec.H - ch) // 2
        ref, srcs, cams, gt = scene.crop_reference(sc, x0, y0, cw, ch)
        e = capi.Engine(ctx, rp, ref, list(srcs), cams)
        e.set_storage(capi.STORE_F16)
        crops.append((ref, srcs, cams, gt, rp))
        cengs.append(e)
    cb = capi.Batch(ctx, cengs)
    cb.set_graph(True)
    cb.prepare()
    init = [(e.depth(), e.rng()) for e in cengs]
    cb.run()
    O.set_threads(min(16, os.cpu_count() or 1))
    f = lambda x: float(np.float32(x))  # noqa: E731
    per_view = []
    for i, (e, (ref, srcs, cams, gt, rp)) in enumerate(zip(cengs, crops)):
        dg = e.depth().astype(np.float64)
        r16, s16 = e.images()
        H = e.homographies()
        res = {"view": i, "size": f"{views[i][0].W}x{views[i][0].H}"}
        for tag, (ri, si) in (("rounded", (r16, s16)), ("fp32", (
                O.blur(ref, rp.gaussian_sigma, rp.blur_mode),
                np.stack([O.blur(s_, rp.gaussian_sigma, rp.blur_mode) for s_ in srcs])))):
            pb = O.Problem(ri, si, H, rp.half_window, f(rp.depth_near), f(rp.depth_far),
                           f(rp.spm_alpha), O.inv_k4(cams[0, :9]), f(rp.orientation_x),
                           f(rp.orientation_z))
            do, _, _ = pb.run(init[i][0].astype(np.float64), init[i][1].copy(), rp.num_samples,
                              rp.num_iterations, record_ids=False)
            res[f"within_1pct_of_oracle_{tag}"] = float(np.mean(np.abs(dg - do) <= 0.01 * np.abs(do)))
            res[f"oracle_{tag}_within_1pct_of_ground_truth"] = float(
                np.mean(np.abs(do - gt) <= 0.01 * gt))
        res["gpu_within_1pct_of_ground_truth"] = float(np.mean(np.abs(dg - gt) <= 0.01 * gt))
        per_view.append(res)
        log(f"view {i} ({res['size']}): {res['within_1pct_of_oracle_rounded']:.4f} within 1% of the "
            f"oracle on the FP16 images, {res['within_1pct_of_oracle_fp32']:.4f} of the oracle on "
            f"FP32 images")
    cb.close()
    for e in cengs:
        e.close()
    ctx.close()
    line["per_view_error"] = {"sample": f"centred {cw}x{ch} crop of each view, batched, FP16",
                              "views": per_view}
    return line


def host_info(threads):
    """The host the CPU baseline ran on: logical CPUs, the CPUs this process
    may use, the CPU model and the oracle threads used."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "usable_cpus": usable, "cpu_model": model,
            "threads": threads}


def cpu_baseline_and_accuracy(sc, rp, sample, ctx, gpu_depth, threads, fp16=False):
    """The FP64 oracle (OpenMP) on the timed view and parameters, from the
    raw images (its own pre-blur, box sums, homographies, generator and
    initial depth), against the engine's depth map of the timed run.
    sample None: the whole view (gpu_depth is the timed run's depth map);
    (w, h): the centred crop, run again on the GPU for the comparison.  With
    fp16 the engine stores FP16 texels while the oracle reads the FP32
    images: the agreement is then the error FP16 storage adds."""
    from oracle import pmx_oracle as O
    from patchmatch_cuda_amd import capi, scene

    W, H = sc.spec.W, sc.spec.H
    if sample is None:
        w, h = W, H
        ref, srcs, cams, gt = sc.ref, sc.srcs, sc.cams, sc.gt_depth
        what = f"the whole {W}x{H} reference view"
    else:
        w, h = sample
        x0, y0 = (W - w) // 2, (H - h) // 2
        ref, srcs, cams, gt = scene.crop_reference(sc, x0, y0, w, h)
        what = f"the centred {w}x{h} crop of the {W}x{H} reference view"
    O.set_threads(threads)
    f = lambda x: float(np.float32(x))  # noqa: E731
    t0 = time.perf_counter()
    ref_b = O.blur(ref, rp.gaussian_sigma, rp.blur_mode)
    srcs_b = np.stack([O.blur(s, rp.gaussian_sigma, rp.blur_mode) for s in srcs])
    pb = O.Problem(ref_b, srcs_b, O.homographies(cams, rp.orientation_x, rp.orientation_z),
                   rp.half_window, f(rp.depth_near), f(rp.depth_far), f(rp.spm_alpha),
                   O.inv_k4(cams[0, :9]), f(rp.orientation_x), f(rp.orientation_z))
    st = O.rng_init_grid(w, h, rp.block_x, rp.block_y, rp.rng_seed)
    d0 = O.init_depth(st, f(rp.depth_near), f(rp.depth_far), rp.init_mode)
    d_o, _, _ = pb.run(d0, st, rp.num_samples, rp.num_iterations, record_ids=False)
    cpu_s = time.perf_counter() - t0
    log(f"cpu oracle on {w}x{h}: {cpu_s:.2f} s with {O.get_threads()} threads")
    hi = host_info(O.get_threads())
    out = {
        "cpu_baseline": {
            "value": (w * h) / cpu_s / 1e6,
            "unit": "Mpix/s",
            "cores": O.get_threads(),
            "kind": "port",
            "sample": (f"FP64 oracle, full pipeline (pre-blur, box sums, RNG, init, "
                       f"{rp.num_iterations} iterations x 4 sweeps) on {what} against all "
                       f"{len(srcs)} full-size sources, OpenMP over sweep columns"),
            "wall_s": cpu_s,
            **hi,
        }
    }
    if sample is None:
        d_g = np.asarray(gpu_depth, np.float64)
    else:
        eng = capi.Engine(ctx, rp, ref, list(srcs), cams)
        if fp16:
            eng.set_storage(capi.STORE_F16)
        eng.prepare()
        eng.run()
        d_g = eng.depth().astype(np.float64)
        eng.close()
    agree = float(np.mean(np.abs(d_g - d_o) <= 0.01 * np.abs(d_o)))
    out["accuracy"] = {
        "within_1pct_of_fp64_oracle": agree,
        "sample": what + (" (the timed run's depth map)" if sample is None else "")
                  + (", FP16 texels against the oracle on FP32 images" if fp16 else ""),
        "pixels": w * h,
        "gpu_within_1pct_of_ground_truth": float(np.mean(np.abs(d_g - gt) <= 0.01 * gt)),
        "oracle_within_1pct_of_ground_truth": float(np.mean(np.abs(d_o - gt) <= 0.01 * gt)),
    }
    log(f"accuracy on {w}x{h}: {agree:.4f} of depths within 1% of the oracle")
    return out, d_o


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2_realistic")
    ap.add_argument("--cpu-sample", default="full",
                    help="'full' (the whole timed view) or a centred crop WxH for the CPU "
                         "oracle baseline and the accuracy gate")
    ap.add_argument("--parity-crop", default="640x480",
                    help="centred crop of the learned path's parity check against the oracle")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (default: min(16, usable CPUs): the CPU share of "
                         "one GPU on the measurement host)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep-waves", type=int, default=None,
                    help="wavefronts per sweep column (1, 2, 4; 0 = automatic; default: the "
                         "engine's)")
    ap.add_argument("--weights", default=None,
                    help="proposal network weights (default: the packaged file)")
    ap.add_argument("--tau", type=float, default=0.2,
                    help="confidence threshold of the learned proposal")
    ap.add_argument("--proposal-sweeps", type=int, default=4,
                    help="sweeps of a learned run that get the proposal (0: all; 4: the "
                         "first iteration)")
    ap.add_argument("--proposal-seeding", type=int, default=0,
                    help="1: the learned proposal also seeds the initial depth")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--fp16", action="store_true", help="FP16 texel storage")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched workload: keep the default launch shape")
    ap.add_argument("--batch-shape", type=int, default=None,
                    help="batched workload: this launch shape (1, 2 or 3), no autotuning")
    ap.add_argument("--batch-crop", default="128x96",
                    help="per-view crop of the batched workload's oracle check")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(a reference view is processed by a single device)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c3_batch":
        print(json.dumps(batch_bench(args)), flush=True)
        return

    from patchmatch_cuda_amd import capi, configs, proposal, scene

    spec, rp = configs.WORKLOADS[args.workload]
    t0 = time.perf_counter()
    sc = scene.render(spec)
    log(f"rendered {args.workload}: {spec.W}x{spec.H}, {spec.V} sources in "
        f"{time.perf_counter() - t0:.1f} s")
    ctx = capi.Context(0)
    eng = capi.Engine(ctx, rp, sc.ref, list(sc.srcs), sc.cams)
    if args.sweep_waves is not None:
        eng.set_sweep_waves(args.sweep_waves)
    if args.fp16:
        eng.set_storage(capi.STORE_F16)
    wall, st = run_gpu_steps(eng, args.steps, args.warmup)
    plain_depth = eng.depth()
    plain_gt = float(np.mean(np.abs(plain_depth - sc.gt_depth) <= 0.01 * sc.gt_depth))
    npx = spec.W * spec.H
    ms_step = wall * 1e3 / args.steps
    value = npx * args.steps / wall / 1e6

    launches = max(st["sweep_launches"], 1)
    avg_ms = st["ms_sweep"] / launches
    flops_per_launch = FLOPS_PER_SAMPLE * st["ncc_samples"] / launches
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    traffic, traffic_src = profiled_traffic(args.workload)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 texels, f32 sums" if args.fp16 else "f32",
        "data": "synthetic (seeded piecewise-planar scene renderer, no dataset)",
        "config": {
            "workload": args.workload,
            "ref": f"{spec.W}x{spec.H}",
            "src_views": spec.V,
            "window": 2 * rp.half_window + 1,
            "iterations": rp.num_iterations,
            "num_samples": rp.num_samples,
            "method": "plain",
            "sweep_waves": eng.sweep_waves(),
        },
        "roofline": {
            "bound": "valu",
            "pipe": "VALU FP32 (NCC warp + bilinear + window sums; no matrix ops)",
            "kernel": "pmx_sweep_kernel",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "avg_launch_ms": avg_ms,
            "samples_per_launch": st["ncc_samples"] / launches,
            "flops_per_sample": FLOPS_PER_SAMPLE,
        },
        "stages_ms_per_step": {
            "sweeps": st["ms_sweep"] / args.steps,
            "cost_all_views": st["ms_cost_all"] / args.steps,
            "transposes": st["ms_transpose"] / args.steps,
            "prepare": st["ms_prepare"] / args.steps,
        },
    }
    line["plain"] = {"value": value, "ms_per_step": ms_step, "sweeps": 4 * rp.num_iterations,
                     "within_1pct_of_ground_truth": plain_gt}
    # host buffers in, depth map out: upload of the reference and sources,
    # device allocation, the step, readback (never the headline value)
    ctx.sync()
    t0 = time.perf_counter()
    e2 = capi.Engine(ctx, rp, sc.ref, list(sc.srcs), sc.cams)
    if args.fp16:
        e2.set_storage(capi.STORE_F16)
    e2.prepare()
    e2.run()
    e2.depth()
    t1 = time.perf_counter()
    e2.close()
    line["pcie_inclusive"] = {"value": npx / (t1 - t0) / 1e6, "unit": "Mpix/s",
                              "ms": (t1 - t0) * 1e3,
                              "what": "engine creation (allocation + upload of 11 float32 "
                                      "images) + one step + depth readback"}
    log(f"PCIe-inclusive: {(t1 - t0) * 1e3:.1f} ms")
    weights = args.weights or proposal.DEFAULT_WEIGHTS
    d_oracle = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, host_info(0)["usable_cpus"])
        sample = None if args.cpu_sample == "full" else parse_sample(args.cpu_sample)
        extra, d_oracle = cpu_baseline_and_accuracy(sc, rp, sample, ctx, plain_depth, threads,
                                                    args.fp16)
        line.update(extra)
    if args.no_learned:
        pass
    elif not os.path.exists(weights):
        line["learned"] = {"skipped": f"no weights at {os.path.relpath(weights, ROOT)}"}
    elif d_oracle is None:
        line["learned"] = {"skipped": "needs the oracle crop (--no-cpu-baseline given)"}
    else:
        lr = learned_leg(ctx, eng, sc, rp, weights, d_oracle,
                         None if args.cpu_sample == "full" else parse_sample(args.cpu_sample),
                         parse_sample(args.parity_crop), args.steps, args.warmup, args.tau,
                         bool(args.proposal_seeding), args.proposal_sweeps)
        # both at the iteration count each needs to reach the accuracy bar
        need = lr["iterations_to_99pct_of_converged_oracle"]
        lr["speedup_vs_plain"] = lr["speedup_vs_plain_at_equal_accuracy"]
        lr["sweep_reduction_vs_plain"] = ((need["plain"] or rp.num_iterations) /
                                          (need["learned"] or rp.num_iterations))
        line["learned"] = lr
    eng.close()
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
