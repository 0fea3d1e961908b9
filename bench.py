#!/usr/bin/env python3
"""Benchmark: Mpoints/s of one full KIFMM evaluation on one MI355X, mixed
precision against FP64 at matched sampled error.

Headline workload: 4,194,304 points i.i.d. uniform in [0,1]^3 (mapped to the
reference root box by x' = 2x - 1), separate targets, densities in [0,1),
seed 1, NP = 4.  A step is one kifmm_gpu_evaluate over inputs already
resident in HBM: up, V, W, X, down and the U-list, and the device -> host
copy of the potentials into host memory (the handle's pinned buffer; the
copies of early U-list chunks overlap the later chunks).  `value` is the
mixed-precision, batched, model-tuned path (the cost model picks points per
box and the precision path, kifmm_gpu_autotune the launch shapes and the
stream layout, one HIP graph per step); `mixed_vs_fp64` holds it beside the
plain FP64 path on the same input, each with its sampled error against the
FP64 direct sum on one shared NV = 1000 sample, and `ratio` is the model-
tuned path over plain FP64.

`plain_fp64_np6` is the plain FP64 drop-in on 4,194,304 uniform points, seed
1, NP = 6, 512 per box: its rate (copy included), the create cost split
into host plan, upload and close-pair search, the one-shot rate (create +
evaluate + copy), a step with fresh densities, and the roofline of its
U-list kernel.  `large_mixed` runs 16,777,216 uniform points (256 per box)
at NP = 3 and 4 on both paths; `tuned_nonuniform` the 1,048,576-point pole-
clustered ellipsoid (NP = 4, 128 per box, FP64, one HIP graph) with default
and autotuned shapes.  Every Mpoints/s includes the copy back.

Prints one JSON line.  Per-kernel HIP event times come from the library's
own streams.  --gpus above 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "kifmm--hybrid--double-only_amd"))

P2P_FLOPS_PER_PAIR = 11.0  # the reference's convention, fmm/driver.cc:60
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector and dense matrix)
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector
HBM_PEAK_GBS = 8000.0
# Instruction-issue ceilings of the U-list kernels' inner loops (gfx950
# assembly of this build; packed FP32 is off, transcendentals are quarter
# rate, and MFMA issue adds to VALU issue rather than overlapping it):
#   mixed: per 1024 pairs 4 v_mfma_f32_16x16x32_bf16 (16 cycles each),
#          16 v_rsq_f32 (16) and 16 v_fmac_f32 (4): 384 cycles per SIMD;
#   FP64 Gram form: per 64 pairs 8 full-rate FP64 instructions (4 cycles each)
#          and one v_rsq_f64 (16): 48 cycles per SIMD.
SIMDS, CLOCK_GHZ = 1024, 2.4
ISSUE_PAIRS_PER_CYCLE = {"mixed": 1024 / 384, "fp64": 64 / 48}


def issue_bound(kind, achieved_tflops):
    t = P2P_FLOPS_PER_PAIR * ISSUE_PAIRS_PER_CYCLE[kind] * SIMDS * CLOCK_GHZ * 1e9 / 1e12
    return {"tflops": t, "frac": achieved_tflops / t, "pairs_per_cycle_per_simd": ISSUE_PAIRS_PER_CYCLE[kind],
            "model": "inner-loop instruction issue at 2.4 GHz on 1024 SIMDs (quarter-rate rsq, no packed FP32)"}
METRIC = "Mpoints/s per FMM eval (4M uniform), mixed vs FP64 GPU at matched sampled rel err"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nv", type=int, default=1000, help="validation samples, direct sums on the GPU (0 = skip)")
    ap.add_argument("--only", choices=("headline", "np6", "nonuniform", "large"), default=None,
                    help="run one case only (profiling): the headline comparison, the NP = 6 FP64 case, "
                         "the autotuned ellipsoid or the 16M-point mixed-vs-FP64 case")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tuned", action="store_true", help="skip the autotuned non-uniform case")
    ap.add_argument("--no-large", action="store_true", help="skip the 16M-point mixed-vs-FP64 case")
    ap.add_argument("--no-gpu-tree", action="store_true", help="skip timing the GPU tree build")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (counter calibration) and exit")
    return ap.parse_args()


def calibrate_copy(mib):
    """A copy kernel of known size: FETCH_SIZE / WRITE_SIZE read on it give
    the scale from counter units to bytes for the traffic figures."""
    import torch

    n = mib * (1 << 20) // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel": "torch.neg float64", "bytes_read": n * 8, "bytes_written": n * 8,
                      "launches": 5}))
    return 0


def read_traffic(name):
    """HBM bytes per launch of a kernel from a committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel)."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    try:
        return float(json.loads(f.read_text())["hbm_bytes_per_launch"])
    except Exception:
        return None


def host_cpu():
    """CPU model, physical cores and the CPUs this process may use."""
    info = {"model": None, "physical_cores": None, "cpus_allowed": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {k.strip(): v.strip() for k, v in (ln.split(":", 1) for ln in out.splitlines() if ":" in ln)}
        info["model"] = kv.get("Model name")
        info["physical_cores"] = int(kv.get("Socket(s)", "1")) * int(kv.get("Core(s) per socket", "0"))
    except Exception:
        pass
    return info


def cpu_baseline(n, dist, q, np_, seed, n_threads):
    """The oracle (restated omp_sse_block, rsqrtps + Newton kernel, OpenMP)
    on the headline input, timed as fmm/evaluate--basic.cc:30-87 times the
    CPU run (operator build included) and evaluate-only."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O

    o = O.Oracle(n, dist, q, np_, seed, 0)
    t0 = time.perf_counter()
    o.build_operators(O.KERNEL_REF)
    t_ops = time.perf_counter() - t0
    _, t = o.evaluate(O.KERNEL_REF, n_threads)
    cpu = host_cpu()
    return {"value": n / t["total"] / 1e6, "unit": "Mpoints/s", "cores": n_threads, "kind": "port",
            "value_with_operators": n / (t["total"] + t_ops) / 1e6, "operators_s": t_ops, "evaluate_s": t["total"],
            "cpu_model": cpu["model"], "host_physical_cores": cpu["physical_cores"],
            "cpus_allowed": cpu["cpus_allowed"],
            "sample": f"one full evaluation of the headline input ({n} uniform points, NP={np_}, {q} pts/box; "
                      f"rsqrtps+Newton kernel, OpenMP over {n_threads} threads); value is evaluate-only, "
                      f"value_with_operators adds the operator build as fmm/evaluate--basic.cc:30-87 times it; "
                      f"phases s: up {t['up']:.2f} U {t['ulist']:.2f} V {t['vlist']:.2f} "
                      f"W {t['wlist']:.2f} X {t['xlist']:.2f} down {t['down']:.2f}"}


def timed(K, prob, precision, steps, warmup, graph=False, tuning=None, copy="view"):
    """Evaluator, mean wall ms per evaluation (inputs resident; copy="view":
    potentials back into the handle's pinned buffer, False: left on the
    device) and mean HIP-event phase times."""
    ev = K.GpuEvaluator(prob, precision=precision, graph=graph)
    if tuning:
        ev.set_tuning(tuning["p2p_form"], tuning["hadamard_fb"], tuning["hadamard_form"], tuning["overlap"],
                      tuning.get("up_form", 0), tuning.get("down_form", 0))
    for _ in range(warmup):
        ev.evaluate(copy=copy)
    phases = []
    t0 = time.perf_counter()
    for _ in range(steps):
        phases.append(ev.evaluate(copy=copy)[1])
    ms = (time.perf_counter() - t0) / steps * 1e3
    avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    return ev, ms, avg


def rate(n, ms):
    return n / (ms * 1e-3) / 1e6


def sampled_error(prob, ev, sample, nv):
    """Sampled relative error on the shared sample (drawn on first use)."""
    if nv <= 0:
        return None, sample
    pot = prob.scatter(ev.evaluate()[0])
    if sample is None:
        err, ids, direct = prob.validate_gpu(ev, pot, nv)
        return err, (ids, direct)
    f = pot[sample[0]]
    return float(np.linalg.norm(sample[1] - f) / np.linalg.norm(f)), sample


def headline(K, steps, warmup, nv):
    """4,194,304 uniform points, seed 1, NP = 4: the plain FP64 path and the
    mixed path at 256 points per box, and the model-tuned path."""
    n = 4194304
    prob = K.Problem.create(n, "uniform", 256, np_=4, seed=1, coincide=0)
    prob.operators()
    out = {"workload": f"uniform {n} points, separate targets, seed 1, NP=4"}
    sample = None
    for name, prec, graph in (("fp64", K.PREC_FP64, False), ("fp64_graph", K.PREC_FP64, True),
                              ("mixed", K.PREC_MIXED, False), ("mixed_graph", K.PREC_MIXED, True)):
        ev, ms, avg = timed(K, prob, prec, steps, warmup, graph)
        err, sample = sampled_error(prob, ev, sample, nv)
        out[name] = {"mpoints_s": rate(n, ms), "ms_per_step": ms, "sampled_rel_err": err, "q": 256}
        if not graph:
            out[name]["phases_ms"] = {k: round(v, 4) for k, v in avg.items()}
        del ev
    # the model-tuned path: the cost model picks q and the precision path for
    # this input (its two best-ranked candidates timed), kifmm_gpu_autotune
    # the launch shapes and the stream layout, one HIP graph per step
    import costmodel as CM

    pts = prob.points()
    pts = tuple(pts[k] for k in ("sx", "sy", "sz", "den", "tx", "ty", "tz"))
    tuner_m = CM.Tuner(CM.CostModel.load())
    t0 = time.perf_counter()
    q, prec, pred, how = tuner_m.choose(pts, 4, verify=2)
    t_choose = time.perf_counter() - t0
    mq, mprec, mpred = tuner_m.model_pick
    mprob = K.Problem.from_points(*pts, pts_max=q, np_=4)
    mprob.operators()
    tuner = K.GpuEvaluator(mprob, precision=prec)
    t0 = time.perf_counter()
    shapes, cand = tuner.autotune(reps=3)
    t_tune = time.perf_counter() - t0
    del tuner
    ev, ms, _ = timed(K, mprob, prec, steps, warmup, graph=True, tuning=shapes)
    err, sample = sampled_error(mprob, ev, sample, nv)
    del ev
    # the same configuration eagerly without the copy: the U-list as one
    # launch, its HIP-event time for the roofline
    ev, ms_nc, avg_nc = timed(K, mprob, prec, steps, warmup, graph=False, tuning=dict(shapes, overlap=0),
                              copy=False)
    wc = ev.work_counts()
    del ev
    out["model_tuned"] = {"mpoints_s": rate(n, ms), "ms_per_step": ms, "sampled_rel_err": err, "q": q,
                          "precision": "mixed" if prec == K.PREC_MIXED else "fp64", "shapes": shapes,
                          "chosen_by": how, "predicted_or_measured_ms": pred, "choose_s": t_choose,
                          "autotune_s": t_tune,
                          "model_only_pick": {"q": mq, "precision": "mixed" if mprec == K.PREC_MIXED else "fp64",
                                              "predicted_ms": mpred},
                          "no_copy": {"ms_per_step": ms_nc, "mpoints_s": rate(n, ms_nc),
                                      "phases_ms": {k: round(v, 4) for k, v in avg_nc.items()}}}
    if (mq, mprec) != (q, prec):  # the network's own pick, timed the same way
        oprob = K.Problem.from_points(*pts, pts_max=mq, np_=4)
        oprob.operators()
        ev, ms_o, _ = timed(K, oprob, mprec, max(3, steps // 2), warmup, graph=True)
        out["model_tuned"]["model_only_pick"]["mpoints_s"] = rate(n, ms_o)
        del ev, oprob
    else:
        out["model_tuned"]["model_only_pick"]["mpoints_s"] = rate(n, ms)
    out["ratio"] = out["model_tuned"]["mpoints_s"] / out["fp64"]["mpoints_s"]
    out["ratio_vs_fp64_graph"] = out["model_tuned"]["mpoints_s"] / out["fp64_graph"]["mpoints_s"]
    out["ratio_mixed_graph_default_q"] = out["mixed_graph"]["mpoints_s"] / out["fp64"]["mpoints_s"]
    return out, wc, avg_nc, prec


def plain_fp64_np6(K, steps, warmup, nv, gpu_tree):
    """4,194,304 uniform points, seed 1, NP = 6, 512 per box, FP64."""
    n = 4194304
    t0 = time.perf_counter()
    prob = K.Problem.create(n, "uniform", 512, np_=6, seed=1, coincide=0)
    prob.operators()
    t_host = time.perf_counter() - t0
    setup = prob.setup_times()
    out = {"workload": f"uniform {n} points, separate targets, seed 1, NP=6, 512 pts/box, fp64"}
    if gpu_tree:
        g = K.Problem.create(n, "uniform", 512, np_=6, seed=1, coincide=0, device=0)
        setup["gpu_tree_lists"] = g.setup_times()["tree_lists"]
        del g
    # one-shot: create + one evaluation with the copy back (the fmmd--hip use)
    c0 = time.perf_counter()
    ev = K.GpuEvaluator(prob)
    c1 = time.perf_counter()
    ev.evaluate(copy="view")
    c2 = time.perf_counter()
    create = ev.create_times()
    wc = ev.work_counts()
    del ev
    ev, ms, avg = timed(K, prob, K.PREC_FP64, steps, warmup)
    err, _ = sampled_error(prob, ev, None, nv)
    # steps with fresh densities: upload (host -> device) + evaluation + copy
    rng = np.random.default_rng(1)
    dens = [rng.random(n) for _ in range(2)]
    for d in dens:
        ev.set_densities(d)
        ev.evaluate(copy="view")
    t0 = time.perf_counter()
    for k in range(max(3, steps // 4)):
        ev.set_densities(dens[k & 1])
        ev.evaluate(copy="view")
    ms_fresh = (time.perf_counter() - t0) / max(3, steps // 4) * 1e3
    # into a caller's pageable array instead of the pinned buffer
    own = np.zeros(n)
    ev.evaluate_into(own)
    t0 = time.perf_counter()
    for _ in range(5):
        ev.evaluate_into(own)
    ms_pageable = (time.perf_counter() - t0) / 5 * 1e3
    del ev
    # the U-list kernel's own time: evaluations without the copy (one launch)
    ev, ms_nc, avg_nc = timed(K, prob, K.PREC_FP64, steps, warmup, copy=False)
    del ev
    p2p_ms = avg_nc["p2p_kernel"]
    achieved = P2P_FLOPS_PER_PAIR * wc["u_pairs"] / (p2p_ms * 1e-3) / 1e12
    out.update({
        "mpoints_s": rate(n, ms), "ms_per_step": ms, "sampled_rel_err": err,
        "phases_ms": {k: round(v, 4) for k, v in avg.items()},
        "no_copy": {"ms_per_step": ms_nc, "mpoints_s": rate(n, ms_nc),
                    "phases_ms": {k: round(v, 4) for k, v in avg_nc.items()}},
        "copy_exposed_ms": avg["copy"], "copy_chunks": wc["copy_chunks"],
        "into_pageable_array": {"ms_per_step": ms_pageable, "mpoints_s": rate(n, ms_pageable)},
        "fresh_densities": {"ms_per_step": ms_fresh, "mpoints_s": rate(n, ms_fresh),
                            "step": "kifmm_gpu_set_densities (host -> device) + evaluation + copy back"},
        "create_ms": create,
        "one_shot": {"create_plus_first_eval_s": c2 - c0, "create_s": c1 - c0, "first_eval_s": c2 - c1,
                     "mpoints_s": n / (c2 - c0) / 1e6,
                     "note": "the process's first evaluation also loads the kernels' code objects"},
        "setup_s": {"host_tree_lists": setup["tree_lists"], "generate": setup["generate"],
                    "operators": setup["operators"], "total_host": t_host,
                    "gpu_tree_lists": setup.get("gpu_tree_lists")},
        "work": {k: wc[k] for k in ("u_pairs", "s2m_pairs", "l2t_pairs", "v_entries", "v_cmac", "gemm_flops",
                                    "close_pairs")},
        "roofline": {"bound": "valu", "pipe": "FP64 VALU (Gram-form pair: 9 instructions, one quarter-rate rsq)",
                     "kernel": "k_p2p_ulist<false, 5, true>", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": read_traffic("p2p_ulist_uniform4M_np6_counters.json"),
                     "algorithmic": f"{P2P_FLOPS_PER_PAIR:g} flop/pair x {wc['u_pairs']:.4g} U-list pairs per launch",
                     "kernel_ms": p2p_ms, "issue_bound": issue_bound("fp64", achieved)}})
    return out


def tuned_nonuniform(K, steps, warmup, nv):
    """1,048,576 ellipsoid points (semi-axes 0.5, 0.5, 0.125, angles drawn
    uniformly: clustered at the poles), seed 2, NP = 4, 128 points per box,
    FP64 in one HIP graph: default launch shapes against the autotuned ones."""
    n = 1048576
    prob = K.Problem.create(n, "ellipsoid", 128, np_=4, seed=2, coincide=0)
    prob.operators()
    out = {"workload": f"ellipsoid {n} points, separate targets, seed 2, NP=4, 128 pts/box, fp64, HIP graph"}
    ev, ms0, _ = timed(K, prob, K.PREC_FP64, steps, warmup, graph=True)
    del ev
    tuner = K.GpuEvaluator(prob)
    t0 = time.perf_counter()
    best, cand = tuner.autotune(reps=3)
    t_tune = time.perf_counter() - t0
    del tuner
    ev, ms1, _ = timed(K, prob, K.PREC_FP64, steps, warmup, graph=True, tuning=best)
    err, _ = sampled_error(prob, ev, None, nv)
    del ev
    ev, _, avg = timed(K, prob, K.PREC_FP64, max(3, steps // 2), warmup, tuning=best)
    del ev
    names = ("p2p_static", "p2p_persistent", "p2p_static64", "hadamard_valu_fb4", "hadamard_valu_fb8",
             "hadamard_valu_fb16", "hadamard_mfma_fb4", "hadamard_mfma_fb8", "hadamard_mfma_fb16",
             "layout0_u_last", "layout1_u_second_stream", "p2p_mixed_split", "layout2_first_chunk_early",
             "layout3_two_chunks_early", "up_gemm", "up_fused", "down_gemm", "down_fused",
             "layout4_u_last_one_stream", "layout5_first_chunk_early_one_stream",
             "layout6_two_chunks_early_one_stream")
    out.update({"default": {"mpoints_s": rate(n, ms0), "ms_per_step": ms0},
                "tuned": {"mpoints_s": rate(n, ms1), "ms_per_step": ms1, "choice": best, "sampled_rel_err": err,
                          "eager_phases_ms": {k: round(v, 4) for k, v in avg.items()}},
                "candidates_ms": {k: round(float(v), 4) for k, v in zip(names, cand)},
                "autotune_s": t_tune})
    return out


def large_mixed(K, steps, warmup, nv):
    """16,777,216 uniform points, seed 3, 256 points per box, at NP = 3 and 4:
    the plain FP64 path and the mixed path (one HIP graph each), Mpoints/s,
    sampled error against the FP64 direct sum and the speedup, with the
    default launch shapes and with each path's autotuned ones.  The tree and
    lists are built on the GPU (identical to the host build)."""
    n = 16777216
    out = {"workload": f"uniform {n} points, separate targets, seed 3, 256 pts/box"}
    for np_ in (3, 4):
        prob = K.Problem.create(n, "uniform", 256, np_=np_, seed=3, coincide=0, device=0)
        prob.operators()
        row, sample = {}, None
        for name, prec in (("fp64", K.PREC_FP64), ("mixed", K.PREC_MIXED)):
            ev, ms, _ = timed(K, prob, prec, steps, warmup, graph=True)
            err, sample = sampled_error(prob, ev, sample, nv)
            row[name] = {"mpoints_s": rate(n, ms), "ms_per_step": ms, "sampled_rel_err": err}
            del ev
            # each path with the launch shapes and layout kifmm_gpu_autotune picks for it
            tuner = K.GpuEvaluator(prob, precision=prec)
            best, _ = tuner.autotune(reps=2)
            del tuner
            ev, ms, _ = timed(K, prob, prec, steps, warmup, graph=True, tuning=best)
            row[name + "_tuned"] = {"mpoints_s": rate(n, ms), "ms_per_step": ms, "shapes": best}
            del ev
        row["speedup"] = row["mixed"]["mpoints_s"] / row["fp64"]["mpoints_s"]
        row["speedup_tuned"] = row["mixed_tuned"]["mpoints_s"] / row["fp64_tuned"]["mpoints_s"]
        out[f"np{np_}"] = row
        del prob
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build evaluates on one GPU; --gpus {args.gpus} is not supported "
              f"(how the evaluation would shard over several GPUs is a written design, not a code path).",
              file=sys.stderr)
        return 2
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    import kifmm_amd as K

    if not K.gpu_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    head = np6 = large = tuned = cpu = None
    roofline = None
    if args.only in (None, "headline"):
        head, wc, avg_nc, prec = headline(K, args.steps, args.warmup, args.nv)
        mt = head["model_tuned"]
        if prec == K.PREC_MIXED:
            kern, peak, pipe = ("k_p2p_ulist_mixed_tau<false>", FP32_PEAK_TFLOPS,
                                "FP32 VALU + bf16 MFMA (per pair: one quarter-rate v_rsq_f32 and one FMA; "
                                "distances on MFMA)")
        else:
            kern, peak, pipe = "k_p2p_ulist<false, 5, true>", FP64_PEAK_TFLOPS, "FP64 VALU"
        p2p_ms = avg_nc["p2p_kernel"]
        achieved = P2P_FLOPS_PER_PAIR * wc["u_pairs"] / (p2p_ms * 1e-3) / 1e12
        roofline = {"bound": "valu", "pipe": pipe, "kernel": kern, "achieved": achieved, "peak": peak,
                    "unit": "TFLOP/s", "frac": achieved / peak,
                    "traffic": read_traffic("p2p_ulist_mixed_uniform4M_np4_counters.json"),
                    "algorithmic": f"{P2P_FLOPS_PER_PAIR:g} flop/pair x {wc['u_pairs']:.4g} U-list pairs per launch",
                    "kernel_ms": p2p_ms, "measured_on": "the model-tuned configuration evaluated eagerly without "
                                                        "the copy (the U-list as one launch)",
                    "issue_bound": issue_bound("mixed" if prec == K.PREC_MIXED else "fp64", achieved)}
    if args.only in (None, "np6"):
        np6 = plain_fp64_np6(K, args.steps, args.warmup, args.nv, not args.no_gpu_tree)
    if args.only in (None, "large") and not args.no_large:
        large = large_mixed(K, max(3, args.steps // 4), 2, args.nv)
    if args.only in (None, "nonuniform") and not args.no_tuned:
        tuned = tuned_nonuniform(K, max(5, args.steps // 2), args.warmup, args.nv)
    if args.only in (None, "headline") and not args.no_cpu_baseline:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
        cpu = cpu_baseline(4194304, "uniform", 256, 4, 1, threads)

    if head is not None:
        mt = head["model_tuned"]
        value, ms_step = mt["mpoints_s"], mt["ms_per_step"]
        config = {"workload": "uniform 4194304 points, separate targets, seed 1, NP=4: the model-tuned path "
                              f"(q={mt['q']}, {mt['precision']}, autotuned shapes, one HIP graph), copy back "
                              "included", "n_points": 4194304, "np": 4, "pts_per_box": mt["q"],
                  "precision": mt["precision"]}
        dtype = "f32+bf16+f64" if mt["precision"] == "mixed" else "f64"
    elif np6 is not None:
        value, ms_step = np6["mpoints_s"], np6["ms_per_step"]
        config = {"workload": np6["workload"] + ", copy back included", "n_points": 4194304, "np": 6,
                  "pts_per_box": 512, "precision": "fp64"}
        dtype = "f64"
        roofline = np6["roofline"]
    elif tuned is not None:
        value, ms_step = tuned["tuned"]["mpoints_s"], tuned["tuned"]["ms_per_step"]
        config = {"workload": tuned["workload"] + ", copy back included", "n_points": 1048576, "np": 4,
                  "pts_per_box": 128, "precision": "fp64"}
        dtype = "f64"
    else:
        value, ms_step = large["np4"]["mixed"]["mpoints_s"], large["np4"]["mixed"]["ms_per_step"]
        config = {"workload": large["workload"] + ", NP=4, mixed, copy back included", "n_points": 16777216,
                  "np": 4, "pts_per_box": 256, "precision": "mixed"}
        dtype = "f32+bf16+f64"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Mpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded drand48 points and densities)",
        "config": config,
        "mixed_vs_fp64": head,
        "plain_fp64_np6": np6,
        "tuned_nonuniform": tuned,
        "large_mixed": large,
        "roofline": roofline,
        "roofline_fp64_np6": np6["roofline"] if np6 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
