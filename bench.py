#!/usr/bin/env python3
"""Benchmark: Mpoints/s of one full KIFMM evaluation on one MI355X.

Workload (default): 4,194,304 points i.i.d. uniform in [0,1]^3 (mapped to
the reference root box by x' = 2x - 1), separate targets, densities in
[0,1), seed 1, NP = 6, 512 points per box -- the plain FP64 GPU case the
reference's fmmd--cuda path runs.  A step is one kifmm_gpu_evaluate over
inputs already resident in HBM (up, U, V, W, X, down); the device -> host
copy of the potentials is reported separately and is not part of `value`.

`mixed_vs_fp64` adds the mixed-precision path against the FP64 path on
4,194,304 uniform points at NP = 4, 256 points per box (same tree), each
with its sampled error against the FP64 direct sum, and the model-tuned
path (cost model's q and precision, autotuned launch shapes, one HIP
graph); `ratio` is that path over the plain FP64 path.  `tuned_nonuniform`
times 1,048,576 pole-clustered ellipsoid points (NP = 4, 128 points per
box, FP64, one HIP graph) with the default launch shapes and with the
shapes kifmm_gpu_autotune picks for that input.  `large_mixed` runs
16,777,216 uniform points (256 per box) at NP = 3 and 4 on both paths.

Prints one JSON line.  Per-kernel HIP event times come from the library's
own stream.  --gpus above 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "kifmm--hybrid--double-only_amd"))

P2P_FLOPS_PER_PAIR = 11.0  # the reference's convention, fmm/driver.cc:60
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector and dense matrix)
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--dist", default="uniform")
    ap.add_argument("--q", type=int, default=512)
    ap.add_argument("--np", dest="np_", type=int, default=6)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--nv", type=int, default=1000, help="validation samples, direct sums on the GPU (0 = skip)")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--no-mixed-compare", action="store_true",
                    help="skip the mixed vs FP64 comparison on 4M uniform points at NP = 4")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tuned", action="store_true", help="skip the autotuned non-uniform case")
    ap.add_argument("--no-gpu-tree", action="store_true", help="skip timing the GPU tree build")
    ap.add_argument("--no-large", action="store_true", help="skip the 16M-point mixed-vs-FP64 case")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-layout-compare", action="store_true",
                    help="skip timing the headline input under the other stream layout")
    ap.add_argument("--overlap", action="store_true",
                    help="time every path with KIFMM_GPU_OVERLAP (the U-list on a second stream alongside up "
                         "and V); by default only the extra 'overlap' figure uses it")
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


def read_traffic():
    """HBM bytes per U-list launch from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel)."""
    f = ROOT / "profiles" / "p2p_ulist_uniform4M_np6_counters.json"
    if not f.exists():
        return None
    try:
        return float(json.loads(f.read_text())["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(args, n_threads):
    """The oracle (restated omp_sse_block, rsqrtps + Newton kernel, OpenMP)
    on the same workload: evaluation timed with operators prebuilt."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O

    o = O.Oracle(args.n, args.dist, args.q, args.np_, args.seed, 0)
    o.build_operators(O.KERNEL_REF)
    _, t = o.evaluate(O.KERNEL_REF, n_threads)
    t_eval = t["total"]
    return {"value": args.n / t_eval / 1e6, "unit": "Mpoints/s", "cores": n_threads, "kind": "port",
            "sample": f"one full evaluation of the same {args.n}-point input (operators prebuilt, "
                      f"rsqrtps+Newton kernel, OpenMP over {n_threads} threads); "
                      f"phases s: up {t['up']:.2f} U {t['ulist']:.2f} V {t['vlist']:.2f} "
                      f"W {t['wlist']:.2f} X {t['xlist']:.2f} down {t['down']:.2f}"}


OVERLAP = False  # KIFMM_GPU_OVERLAP for every timed evaluator (bench.py --overlap)


def timed(K, prob, precision, steps, warmup, graph=False, overlap=None):
    """Mean wall ms per evaluation (inputs resident, result left on the
    device) and mean HIP-event phase times."""
    ev = K.GpuEvaluator(prob, precision=precision, graph=graph, overlap=OVERLAP if overlap is None else overlap)
    for _ in range(warmup):
        ev.evaluate(copy=False)
    phases = []
    t0 = time.perf_counter()
    for _ in range(steps):
        phases.append(ev.evaluate(copy=False)[1])
    ms = (time.perf_counter() - t0) / steps * 1e3
    avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    return ev, ms, avg


def mixed_vs_fp64(K, steps, warmup, nv):
    """4,194,304 uniform points, seed 1, NP = 4, 256 points per box: the
    plain FP64 path and the mixed path on the same tree, Mpoints/s and
    sampled error of each."""
    n = 4194304
    prob = K.Problem.create(n, "uniform", 256, np_=4, seed=1, coincide=0)
    prob.operators()
    out = {"workload": f"uniform {n} points, separate targets, seed 1, NP=4, 256 pts/box"}
    sample = None  # one NV-point sample for every path (matched error)
    for name, prec, graph in (("fp64", K.PREC_FP64, False), ("fp64_graph", K.PREC_FP64, True),
                              ("mixed", K.PREC_MIXED, False), ("mixed_graph", K.PREC_MIXED, True)):
        ev, ms, avg = timed(K, prob, prec, steps, warmup, graph)
        err = None
        if nv > 0:
            pot = prob.scatter(ev.evaluate(copy=True)[0])
            if sample is None:
                err, ids, direct = prob.validate_gpu(ev, pot, nv)
                sample = (ids, direct)
            else:
                f = pot[sample[0]]
                err = float(np.linalg.norm(sample[1] - f) / np.linalg.norm(f))
        out[name] = {"mpoints_s": n / (ms * 1e-3) / 1e6, "ms_per_step": ms, "sampled_rel_err": err}
        if not graph:
            out[name]["phases_ms"] = {k: round(v, 4) for k, v in avg.items()}
        del ev
    # the model-tuned path: the cost model picks q and the precision path for
    # this input, kifmm_gpu_autotune the launch shapes, one HIP graph per step
    import costmodel as CM
    pts = prob.points()
    pts = tuple(pts[k] for k in ("sx", "sy", "sz", "den", "tx", "ty", "tz"))
    t0 = time.perf_counter()
    q, prec, pred, how = CM.Tuner(CM.CostModel.load()).choose(pts, 4, verify=2)
    t_choose = time.perf_counter() - t0
    mprob = K.Problem.from_points(*pts, pts_max=q, np_=4)
    mprob.operators()
    tuner = K.GpuEvaluator(mprob, precision=prec)
    shapes, _ = tuner.autotune(reps=3)
    del tuner
    ev = K.GpuEvaluator(mprob, precision=prec, graph=True, overlap=OVERLAP)
    ev.set_tuning(shapes["p2p_form"], shapes["hadamard_fb"], shapes["hadamard_form"], shapes["overlap"])
    for _ in range(warmup):
        ev.evaluate(copy=False)
    t0 = time.perf_counter()
    for _ in range(steps):
        ev.evaluate(copy=False)
    ms = (time.perf_counter() - t0) / steps * 1e3
    err = None
    if nv > 0 and sample is not None:
        f = mprob.scatter(ev.evaluate(copy=True)[0])[sample[0]]
        err = float(np.linalg.norm(sample[1] - f) / np.linalg.norm(f))
    out["model_tuned"] = {"mpoints_s": n / (ms * 1e-3) / 1e6, "ms_per_step": ms, "sampled_rel_err": err,
                          "q": q, "precision": "mixed" if prec == K.PREC_MIXED else "fp64", "shapes": shapes,
                          "chosen_by": how, "predicted_ms": pred, "choose_s": t_choose}
    del ev
    # the mixed-precision, batched, model-tuned path against the plain FP64 path
    out["ratio"] = out["model_tuned"]["mpoints_s"] / out["fp64"]["mpoints_s"]
    out["ratio_mixed_graph_default_q"] = out["mixed_graph"]["mpoints_s"] / out["fp64"]["mpoints_s"]
    out["ratio_both_graphed"] = out["mixed_graph"]["mpoints_s"] / out["fp64_graph"]["mpoints_s"]
    return out


def tuned_nonuniform(K, steps, warmup, nv):
    """1,048,576 ellipsoid points (semi-axes 0.5, 0.5, 0.125, angles drawn
    uniformly: clustered at the poles), seed 2, NP = 4, 128 points per box,
    FP64 in one HIP graph: default launch shapes against the autotuned ones."""
    n = 1048576
    prob = K.Problem.create(n, "ellipsoid", 128, np_=4, seed=2, coincide=0)
    prob.operators()
    out = {"workload": f"ellipsoid {n} points, separate targets, seed 2, NP=4, 128 pts/box, fp64, HIP graph"}
    _, ms0, _ = timed(K, prob, K.PREC_FP64, steps, warmup, graph=True)
    tuner = K.GpuEvaluator(prob)
    t0 = time.perf_counter()
    best, cand = tuner.autotune(reps=3)
    t_tune = time.perf_counter() - t0
    del tuner
    ev = K.GpuEvaluator(prob, graph=True, overlap=OVERLAP)
    ev.set_tuning(best["p2p_form"], best["hadamard_fb"], best["hadamard_form"], best["overlap"])
    for _ in range(warmup):
        ev.evaluate(copy=False)
    t0 = time.perf_counter()
    for _ in range(steps):
        ev.evaluate(copy=False)
    ms1 = (time.perf_counter() - t0) / steps * 1e3
    err = None
    if nv > 0:
        err, _, _ = prob.validate_gpu(ev, prob.scatter(ev.evaluate(copy=True)[0]), nv)
    names = ("p2p_static", "p2p_persistent", "p2p_static64", "hadamard_valu_fb4", "hadamard_valu_fb8",
             "hadamard_valu_fb16", "hadamard_mfma_fb4", "hadamard_mfma_fb8", "hadamard_mfma_fb16",
             "layout_one_stream", "layout_overlap")
    out.update({"default": {"mpoints_s": n / (ms0 * 1e-3) / 1e6, "ms_per_step": ms0},
                "tuned": {"mpoints_s": n / (ms1 * 1e-3) / 1e6, "ms_per_step": ms1, "choice": best,
                          "sampled_rel_err": err},
                "candidates_ms": {k: round(float(v), 4) for k, v in zip(names, cand)},
                "autotune_s": t_tune})
    return out


def large_mixed(K, steps, warmup, nv):
    """16,777,216 uniform points, seed 3, 256 points per box, at NP = 3 and 4:
    the plain FP64 path and the mixed path (one HIP graph each), Mpoints/s,
    sampled error against the FP64 direct sum and the speedup.  The tree and
    lists are built on the GPU (identical to the host build)."""
    n = 16777216
    out = {"workload": f"uniform {n} points, separate targets, seed 3, 256 pts/box"}
    for np_ in (3, 4):
        prob = K.Problem.create(n, "uniform", 256, np_=np_, seed=3, coincide=0, device=0)
        prob.operators()
        row, sample = {}, None
        for name, prec in (("fp64", K.PREC_FP64), ("mixed", K.PREC_MIXED)):
            ev, ms, _ = timed(K, prob, prec, steps, warmup, graph=True)
            err = None
            if nv > 0:
                pot = prob.scatter(ev.evaluate(copy=True)[0])
                if sample is None:
                    err, ids, direct = prob.validate_gpu(ev, pot, nv)
                    sample = (ids, direct)
                else:
                    f = pot[sample[0]]
                    err = float(np.linalg.norm(sample[1] - f) / np.linalg.norm(f))
            row[name] = {"mpoints_s": n / (ms * 1e-3) / 1e6, "ms_per_step": ms, "sampled_rel_err": err}
            del ev
        row["speedup"] = row["mixed"]["mpoints_s"] / row["fp64"]["mpoints_s"]
        out[f"np{np_}"] = row
        del prob
    return out


def main():
    global OVERLAP
    args = parse()
    OVERLAP = args.overlap
    if args.gpus != 1:
        print(f"bench.py: this build evaluates on one GPU; --gpus {args.gpus} is not supported "
              f"(how the evaluation would shard over several GPUs is a written design, not a code path).", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    import kifmm_amd as K

    if not K.gpu_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    t0 = time.perf_counter()
    prob = K.Problem.create(args.n, args.dist, args.q, np_=args.np_, seed=args.seed, coincide=0)
    prob.operators()
    t_setup = time.perf_counter() - t0
    setup = prob.setup_times()
    setup["gpu_tree_lists"] = None
    if not args.no_gpu_tree:
        gprob = K.Problem.create(args.n, args.dist, args.q, np_=args.np_, seed=args.seed, coincide=0, device=0)
        setup["gpu_tree_lists"] = gprob.setup_times()["tree_lists"]
        del gprob
    prec = K.PREC_MIXED if args.precision == "mixed" else K.PREC_FP64
    t0 = time.perf_counter()
    ev = K.GpuEvaluator(prob, precision=prec)
    t_create = time.perf_counter() - t0
    wc = ev.work_counts()
    del ev
    ev, ms_step, avg = timed(K, prob, prec, args.steps, args.warmup)
    value = args.n / (ms_step * 1e-3) / 1e6

    # one step with the copy back, for the PCIe-inclusive figure and validation
    pot_leaf, tc = ev.evaluate(copy=True)
    pot = prob.scatter(pot_leaf)
    err = None
    if args.nv > 0:
        err, _, _ = prob.validate_gpu(ev, pot, args.nv)

    p2p_ms = avg["p2p_kernel"]
    achieved = P2P_FLOPS_PER_PAIR * wc["u_pairs"] / (p2p_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "pipe": "FP64 VALU (MI355X FP64 vector peak equals its dense FP64 MFMA peak)",
                "kernel": "k_p2p_ulist", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": read_traffic(),
                "algorithmic": f"{P2P_FLOPS_PER_PAIR:g} flop/pair x {wc['u_pairs']:.4g} U-list pairs per launch",
                "kernel_ms": p2p_ms,
                "kernel_shares_gpu": OVERLAP}
    del ev
    # the same input with the other stream layout: KIFMM_GPU_OVERLAP runs the
    # U-list alongside up and V (faster overall; the U-list kernel then
    # shares the chip and its own span grows)
    other = None
    if not args.no_layout_compare:
        ev_o, ms_o, avg_o = timed(K, prob, prec, args.steps, args.warmup, overlap=not OVERLAP)
        del ev_o
        other = {"overlap": not OVERLAP, "ms_per_step": ms_o, "mpoints_s": args.n / (ms_o * 1e-3) / 1e6,
                 "p2p_kernel_ms": avg_o["p2p_kernel"]}

    mixed = None
    if not args.no_mixed_compare:
        mixed = mixed_vs_fp64(K, max(5, args.steps // 2), args.warmup, args.nv)

    large = None
    if not args.no_large:
        large = large_mixed(K, max(3, args.steps // 4), 2, args.nv)

    tuned = None
    if not args.no_tuned:
        tuned = tuned_nonuniform(K, max(5, args.steps // 2), args.warmup, args.nv)

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        cpu = cpu_baseline(args, threads)

    out = {
        "metric": "Mpoints/s per FMM eval (4M uniform), mixed vs FP64 GPU at matched sampled rel err",
        "value": value,
        "unit": "Mpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded drand48 points and densities)",
        "config": {"workload": f"{args.dist} {args.n} points, separate targets, seed {args.seed}, "
                               f"NP={args.np_}, {args.q} pts/box, {args.precision}",
                   "n_points": args.n, "np": args.np_, "pts_per_box": args.q, "precision": args.precision},
        "phases_ms": {k: round(v, 4) for k, v in avg.items()},
        "mpoints_s_incl_d2h": args.n / ((ms_step + tc["copy"]) * 1e-3) / 1e6,
        "d2h_ms": tc["copy"],
        "sampled_rel_err": err,
        "setup_s": {"host_tree_lists": setup["tree_lists"], "generate": setup["generate"],
                    "operators": setup["operators"], "upload": t_create, "total_host": t_setup,
                    "gpu_tree_lists": setup["gpu_tree_lists"]},
        "work": {k: wc[k] for k in ("u_pairs", "s2m_pairs", "l2t_pairs", "v_entries", "v_cmac", "gemm_flops")},
        "mixed_vs_fp64": mixed,
        "tuned_nonuniform": tuned,
        "large_mixed": large,
        "roofline": roofline,
        "stream_layout": {"overlap": OVERLAP, "other_layout": other},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
