#!/usr/bin/env python
"""PO far-field throughput on one MI355X: tree solve vs the FP64 direct sum.

Metric: point-direction pairs/s = N_points x N_output_directions x S_solves
per second of the batched tree solve, gated on <= 1e-3 relative L2 of the
intensity against the FP64 depth-0 direct sum of the same points: the HIP
FP64 direct sum over the full grid (every solve), and the C FP64 oracle over
the sample of output directions that the CPU baseline evaluates.  A failed
gate prints value null and exits 1; a skipped gate (--no-direct) prints
value null.

Default workload (one GPU): the synthetic fiber segment -- elliptical cross
section 40 x 30 um, length 30 um, sawtooth cuticle scales (period 5 um, tilt
3 deg), Gaussian height noise sigma 0.05 um / correlation 0.5 um (seed 0),
0.1 um sample spacing, N = 663,000 points, IOR 1.55 + 0.02i, octree depth 6,
32 incident directions (theta index 60, 75, 90, 105 x phi index 0, 45, ..,
315), TM, lambda 0.4 um, far field 360 (theta) x 720 (phi).  A step is one
batched solve of all 32 incidences with the octree of the geometry cached
(inputs resident in HBM).

The tree workloads run at the library's defaults (opts.beta = 4) and call
wf_autotune before the warm-up (--no-autotune skips it): it picks the
transition kernels and the level sampling constant, lowering beta to 2 when
the solves stay within 1e-4 relative L2 of the beta = 4 solves (on the sphere
and fiber workloads they stay within ~2e-5 of the FP64 direct sum, measured
in the same run, and the levels are ~20 % cheaper).  Gradients of
log-intensity losses weight dim directions by 1/I and keep opts.beta (the
adjoint workload).

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6


def workload(name: str):
    import wavefiber3d_amd as wf

    if name == "sphere10um_500k":
        lam = [wf.reference_lambda(l) for l in range(25)]
        incs = [wf.Incidence(math.pi / 2, 0.0, L, p, 1.55 + 0.1j) for L in lam for p in "ME"]
        return dict(surface=lambda: wf.fibonacci_sphere(500_000, 10e-6), depth=4, incs=incs,
                    grid=(180, 360),
                    desc="Fibonacci sphere r=10um N=500000, IOR 1.55+0.1i, depth 4, 25 lambdas "
                         "400-688 nm, +z incidence TM+TE, far field 180x360")
    if name == "sphere2um_12k":
        incs = [wf.Incidence(math.pi / 2, 0.0, 0.5e-6, p, 1.55 + 0j) for p in "ME"]
        return dict(surface=lambda: wf.fibonacci_sphere(12_000, 2e-6), depth=2, incs=incs,
                    grid=(90, 180),
                    desc="Fibonacci sphere r=2um N=12000, IOR 1.55, lambda 0.5um, TM+TE, 90x180")
    if name == "fiber_663k":
        incs = [wf.Incidence(wf.reference_incidence_theta(t), wf.reference_incidence_phi(p),
                             0.4e-6, "M", 1.55 + 0.02j)
                for t in (60, 75, 90, 105) for p in range(0, 360, 45)]
        return dict(surface=lambda: wf.fiber_segment(40e-6, 30e-6, 30e-6, 0.1e-6, seed=0),
                    depth=6, incs=incs, grid=(360, 720),
                    desc="fiber segment 40x30um ellipse, length 30um, cuticle 5um/3deg, noise "
                         "0.05um/0.5um seed 0, 0.1um spacing, IOR 1.55+0.02i, depth 6, "
                         "32 incidences TM, lambda 0.4um, far field 360x720")
    if name == "fibers256":
        # 8 incidences (theta idx {60, 90, 105, 120} x phi idx {0, 180}) x 25 wavelengths
        incs = [wf.Incidence(wf.reference_incidence_theta(t), wf.reference_incidence_phi(p),
                             wf.reference_lambda(l), "M", 1.55 + 0.02j)
                for l in range(25) for t in (60, 90, 105, 120) for p in (0, 180)]
        return dict(batch=True, incs=incs, grid=(360, 720), depth="auto",
                    desc="uneven fiber segments (seeds 0..K-1, r1 U[20,50]um, r2 0.75 r1, length "
                         "U[5,40]um, 20k-400k points), 8 incidences x 25 lambdas TM, "
                         "IOR 1.55+0.02i, far field 360x720, auto depth, batched on worker streams")
    if name == "adjoint_fiber200k":
        return dict(adjoint=True, incs=[wf.Incidence(wf.reference_incidence_theta(90), 0.0, 0.4e-6, "M",
                                                     1.55 + 0.02j)],
                    grid=(180, 360), depth="auto",
                    desc="fiber 40x30um ellipse, length 10um (100 rings), arcs 0..1999 of 2210 "
                         "(N=200000), cuticle 5um/3deg, noise seed 0, IOR 1.55+0.02i, theta idx 90 "
                         "TM, lambda 0.4um, 180x360; loss = sum (log I - log T)^2, T = I_fp64 "
                         "x exp(0.3 N(0,1)) seed 12345")
    if name == "heightfield_fiber400um":
        return dict(heightfield=True,
                    desc="fiber heightfield 40x30um ellipse, length 400um, 0.1um spacing, "
                         "seed 0 (2210 arc x 4000 rings = 8.84M samples) -> points, normals, "
                         "areas (util/geometry.h:78-178)")
    if name == "average_25x7200x1440":
        return dict(average=True,
                    desc="block average of 25 intensity grids 7200x1440 -> 180x360 "
                         "(util/fields.h:206-234)")
    if name == "finite_ellipsoid_ring":
        return dict(finite=True,
                    desc="finite-distance FP64 validator: ellipsoid 3x3x2um mesh (30322 samples), "
                         "500 nm, IOR 1.55+0.3i, incidence +x TM, beam sigma 1 cm, D = 1 m, "
                         "grid 2x3600 (the theta = 0 ring and the south pole ring) "
                         "(compareBEM3d/ellipsoid_simple.cu:384-589)")
    if name == "rough2d_sweep":
        return dict(rough2d=True,
                    desc="2D cross-section PO: reference ellipse (2000 samples), 25 wavelengths "
                         "400-688 nm x TM+TE, incidence phi = pi, IOR 1.55+0.1i, 3600 azimuths, "
                         "FP64 (compareBEM2d/rough2d.cu:393-583)")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(surf, incs, grid, budget_s: float):
    """FP64 oracle direct sum on the host cores over a sample of output
    directions (every k-th), both solves of the first wavelength.  The sample
    grows geometrically until one run takes about budget_s."""
    import oracle

    dirs_all = oracle.output_dirs(*grid)
    lam0 = incs[0].lambda_m
    sample_incs = [t.as_tuple() for t in incs if t.lambda_m == lam0][:2]
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    arrays = surf.arrays()

    def run(nd):
        # every step-th direction over the whole grid (no truncation: the
        # sample spans every row, the intensity peak included)
        step = max(1, -(-len(dirs_all) // nd))
        dirs = dirs_all[::step]
        t0 = time.perf_counter()
        I, _ = oracle.solve_direct(arrays, sample_incs, 0, 0, dirs=dirs, nthreads=nthreads)
        return dirs, step, time.perf_counter() - t0, I

    nd = min(len(dirs_all), 4 * nthreads)
    dirs, step, dt, I = run(nd)
    while dt < 0.6 * budget_s and nd < len(dirs_all):
        grow = budget_s / max(dt, 1e-3)
        nd = min(len(dirs_all), max(nd + 1, int(nd * min(grow, 8.0))))
        dirs, step, dt, I = run(nd)
    pairs = surf.n * len(dirs) * len(sample_incs)
    base = {"value": pairs / dt, "unit": "pairs/s", "cores": nthreads, "kind": "port",
            "sample": f"FP64 oracle direct sum (C, OpenMP), {len(dirs)} of {len(dirs_all)} output "
                      f"directions (every {step}th), {len(sample_incs)} solves at lambda "
                      f"{lam0 * 1e9:.0f} nm, N={surf.n}, {dt:.1f} s"}
    # the oracle's intensities on the sample: solve indices and flat direction
    # indices, for the accuracy gate against the C oracle
    solve_idx = [k for k, t in enumerate(incs) if t.lambda_m == lam0][:2]
    return base, {"solves": solve_idx, "dir_index": np.arange(len(dirs)) * step, "I": I}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fiber_663k")
    ap.add_argument("--precision", default="bf16x3", choices=["fp32", "bf16x3", "fp32_valu"])
    ap.add_argument("--beta", type=float, default=4.0,
                    help="opts.beta, the level sampling constant (the library default 4; "
                         "wf_autotune may lower it for the forward solves under its 1e-4 gate)")
    ap.add_argument("--no-direct", action="store_true", help="skip the FP64 direct-sum gate")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--depth", default=None, help="override the workload's depth (int or 'auto')")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip wf_autotune before the warm-up (tree workloads; on by default)")
    ap.add_argument("--agg-fft-rows", type=int, default=None,
                    help="tune.agg_fft_rows: levels with 2 nt above it use the FFT transition")
    ap.add_argument("--agg-fused", type=int, default=None, choices=[0, 1, 2],
                    help="tune.agg_fused: one-kernel internal level transitions (2, 1) or two passes")
    ap.add_argument("--theta-tile", type=int, default=None, choices=[0, 1, 2, 3],
                    help="tune.theta_tile: block tile of the two-pass theta pass of 6-9 K steps")
    ap.add_argument("--phi-rows", type=int, default=None, choices=[0, 16, 32],
                    help="tune.phi_rows: rows per staged tile of the two-pass phi pass of long rows")
    ap.add_argument("--chunk-solves", type=int, default=None,
                    help="tune.chunk_solves: solves per tree pass (1..8)")
    ap.add_argument("--leaf-quad", type=int, default=None, choices=[0, 1],
                    help="tune.leaf_quad: quad-symmetric leaf kernel where eligible")
    ap.add_argument("--segments", type=int, default=32, help="fibers256: segments in the batch")
    ap.add_argument("--check-all", action="store_true",
                    help="fibers256: check every segment's first solve against the FP64 direct sum")
    ap.add_argument("--streams", type=int, default=4, help="fibers256: worker streams")
    ap.add_argument("--graph", action="store_true", help="capture solves into hipGraphs")
    ap.add_argument("--solve-streams", type=int, default=0,
                    help="opts.solve_streams: concurrent chunks of a solve (0 = library default 4)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one 1 GiB device-to-device copy (reference for the PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent solves shard across GPUs with no exchange; see the design notes)",
              file=sys.stderr)
        return 2

    import torch

    if args.copy_calibration:
        n = 1 << 28  # 1 GiB of float32
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device-to-device copy", "bytes": 4 * n}), flush=True)
        return 0

    import wavefiber3d_amd as wf

    W = workload(args.workload)
    if args.depth is not None:
        W["depth"] = "auto" if args.depth == "auto" else int(args.depth)
    print(f"[bench] workload {args.workload}: {W['desc']}", file=sys.stderr, flush=True)
    if W.get("batch"):
        return bench_batch(args, W)
    if W.get("heightfield"):
        return bench_heightfield(args, W)
    if W.get("average"):
        return bench_average(args, W)
    if W.get("adjoint"):
        return bench_adjoint(args, W)
    if W.get("finite"):
        return bench_finite(args, W)
    if W.get("rough2d"):
        return bench_rough2d(args, W)
    surf = W["surface"]()
    incs = W["incs"]
    nt, np_ = W["grid"]
    D = nt * np_
    S = len(incs)
    dev = [torch.from_numpy(a).cuda() for a in surf.arrays()]
    solver = wf.POSolver(device=0, precision=args.precision, beta=args.beta, use_graph=args.graph,
                         solve_streams=args.solve_streams)
    if args.agg_fft_rows is not None:
        solver.tune = {"agg_fft_rows": args.agg_fft_rows}
    if args.agg_fused is not None:
        solver.tune = {"agg_fused": args.agg_fused}
    if args.theta_tile is not None:
        solver.tune = {"theta_tile": args.theta_tile}
    if args.phi_rows is not None:
        solver.tune = {"phi_rows": args.phi_rows}
    if args.chunk_solves is not None:
        solver.tune = {"chunk_solves": args.chunk_solves}
    if args.leaf_quad is not None:
        solver.tune = {"leaf_quad": args.leaf_quad}
    t0 = time.perf_counter()
    solver.set_geometry(dev, W["depth"])
    torch.cuda.synchronize()
    tree_ms = (time.perf_counter() - t0) * 1e3
    print(f"[bench] geometry + octree {tree_ms:.1f} ms (N={surf.n})", file=sys.stderr, flush=True)
    tuned = None
    if args.autotune:
        t0 = time.perf_counter()
        best = solver.autotune(incs, (nt, np_))
        tuned = {"best_depth": best, "tune": solver.tune,
                 "ms": (time.perf_counter() - t0) * 1e3}
        print(f"[bench] autotune {tuned}", file=sys.stderr, flush=True)
    tb = solver.tune["beta"]
    beta_eff = tb if tb > 0 else args.beta  # the sampling constant the solves use
    stream = torch.cuda.Stream() if args.graph else torch.cuda.current_stream()
    torch.cuda.set_stream(stream)

    for _ in range(args.warmup):
        I = solver.solve(incs, (nt, np_), out="torch")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        I = solver.solve(incs, (nt, np_), out="torch")
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / max(1, args.steps)
    st = solver.stats()
    pairs = float(surf.n) * D * S
    value = pairs / (ms_step * 1e-3)
    print(f"[bench] tree: {ms_step:.2f} ms/step, {value:.3e} pairs/s; kernels "
          + ", ".join(f"{k}={v['ms']:.2f}ms/{v['launches']}" for k, v in st["kernels"].items()),
          file=sys.stderr, flush=True)

    # Roofline.  The timed solves run their chunks on several streams at once,
    # so their per-launch durations include the other streams' kernels beside
    # them; the roofline takes the durations (HIP events around each launch on
    # its stream) of the same solve run sequentially (solve_streams = 1) right
    # after the timed region.
    roof_src = st
    concurrent = not args.graph and args.solve_streams != 1
    if concurrent:
        with wf.POSolver(device=0, precision=args.precision, beta=args.beta, solve_streams=1) as sq:
            sq.tune = solver.tune
            sq.set_geometry(dev, W["depth"])
            for _ in range(2):
                sq.solve(incs, (nt, np_), out="torch")
            torch.cuda.synchronize()
            roof_src = sq.stats()
            seq_ms = roof_src["ms_total"]
    kern = roof_src["kernels"]
    per_kernel = {}
    for k, v in kern.items():
        if not v["launches"]:
            continue
        avg = v["ms"] / v["launches"]
        e = {"launches": v["launches"], "avg_launch_ms": avg, "work_per_launch": v["work"] / v["launches"]}
        if k in ("leaf", "direct_fp64"):
            e["tflops"] = e["work_per_launch"] / (avg * 1e-3) / 1e12
        elif v["work"] > 0:
            e["gbs"] = e["work_per_launch"] / (avg * 1e-3) / 1e9
        per_kernel[k] = e
    # aggregation: one "launch" = one level transition (phi pass + theta pass
    # + translation table), work = compulsory bytes (children read once,
    # parents written once), time = the summed durations of its kernels
    agg_ms = sum(kern[k]["ms"] for k in ("phi_pass", "theta_pass", "trans_table") if k in kern)
    leaf_ms = kern.get("leaf", {"ms": 0.0})["ms"]
    if roof_src.get("agg_transitions", 0) and agg_ms >= leaf_ms:
        ntr = roof_src["agg_transitions"]
        avg_ms = agg_ms / ntr
        work = roof_src["agg_bytes"] / ntr
        achieved = work / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": "aggregation (level transition: phi + theta + table)",
                "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "work_per_launch": work, "avg_launch_ms": avg_ms, "launches": ntr,
                "work_definition": "compulsory HBM bytes per level transition: 8 B per child "
                                   "pattern sample and field read once + 8 B per parent sample "
                                   "and field written once (16 B at the FP64 root)"}
    else:
        kd = kern["leaf"]
        avg_ms = kd["ms"] / kd["launches"]
        work = kd["work"] / kd["launches"]
        achieved = work / (avg_ms * 1e-3) / 1e12
        peak = BF16_PEAK_TFLOPS if args.precision == "bf16x3" else FP32_PEAK_TFLOPS
        roof = {"bound": "mfma", "kernel": "leaf", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "work_per_launch": work, "avg_launch_ms": avg_ms, "launches": kd["launches"],
                "work_definition": "useful FLOP: 8 per complex multiply-add of the leaf sum"}
    roof["per_kernel"] = per_kernel
    if concurrent:
        roof["measured"] = (f"sequential pass (solve_streams=1, {seq_ms:.2f} ms per solve), "
                            "HIP events on the solve stream")
    # HBM bytes per level transition measured with rocprofv3 FETCH_SIZE /
    # WRITE_SIZE on this workload (calibrated on a device copy), kept beside
    # the profiles
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         f"traffic_{args.workload}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if (tr.get("kernel") == roof["kernel"]
                and tr.get("precision", args.precision) == args.precision
                and tr.get("beta", 4.0) == beta_eff):
            roof["traffic"] = tr["traffic_bytes_per_launch"]
            roof["traffic_source"] = os.path.relpath(tpath, os.path.dirname(os.path.abspath(__file__)))

    out = {
        "metric": "PO point-direction pairs/s (tree vs direct) at \u22641e-3 rel-L2 vs FP64 oracle",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "bf16x3-f32acc",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": W["desc"], "n_points": surf.n,
                   "n_directions": D, "solves": S, "depth": W["depth"],
                   "depth_used": st["depth"], "precision": args.precision, "beta": beta_eff,
                   "beta_opts": args.beta,
                   "graph": args.graph, "solve_streams": args.solve_streams or 4,
                   "autotune": tuned, "tune": solver.tune,
                   "levels": st["levels"]},
        "roofline": roof,
        "solves_per_s": S / (ms_step * 1e-3),
        "kernels_ms": {k: v["ms"] for k, v in kern.items()},
        "kernels_ms_note": ("summed launch durations of one solve run sequentially; the timed "
                            "solves overlap chunks on several streams" if concurrent else
                            "summed launch durations of the last timed solve"),
        "geometry_octree_ms": tree_ms,
    }

    if not args.no_direct:
        I_tree = I.cpu().numpy().reshape(S, -1).astype(np.float64)
        t0 = time.perf_counter()
        Id = solver.solve_fp64_direct(dev, incs, (nt, np_))
        torch.cuda.synchronize()
        dms = (time.perf_counter() - t0) * 1e3
        std = solver.stats()
        Id = Id.cpu().numpy().reshape(S, -1)
        err = np.linalg.norm(I_tree - Id, axis=1) / np.linalg.norm(Id, axis=1)
        lam = np.array([t.lambda_m for t in incs])
        per_lam = {f"{L * 1e9:.0f}nm": float(err[lam == L].max()) for L in sorted(set(lam))}
        dk = std["kernels"].get("direct_fp64", {"ms": dms, "launches": 1, "work": 48.0 * pairs})
        out["direct_fp64"] = {
            "ms": dms, "pairs_per_s": pairs / (dms * 1e-3),
            "kernel_ms": dk["ms"],
            "kernel_tflops": dk["work"] / (dk["ms"] * 1e-3) / 1e12,
            "frac_fp64_peak": dk["work"] / (dk["ms"] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        }
        out["speedup_vs_fp64_direct"] = out["direct_fp64"]["ms"] / ms_step
        out["accuracy"] = {"rel_l2_max": float(err.max()), "rel_l2_per_lambda": per_lam,
                           "reference": "HIP FP64 direct sum, full grid, every solve (it matches "
                                        "the C FP64 oracle to <= 1e-12 in the GPU tests)",
                           "gate": 1e-3}
        print(f"[bench] FP64 direct {dms:.1f} ms; speedup {out['speedup_vs_fp64_direct']:.1f}x; "
              f"rel-L2 max {err.max():.2e}", file=sys.stderr, flush=True)

    if not args.no_cpu:
        out["cpu_baseline"], smp = cpu_baseline(surf, incs, (nt, np_), args.cpu_budget)
        print(f"[bench] cpu baseline {out['cpu_baseline']['value']:.3e} pairs/s", file=sys.stderr,
              flush=True)
        # the same sample against the tree: the gate on the C FP64 oracle itself
        It = I.cpu().numpy().reshape(S, -1).astype(np.float64)
        eo = [float(np.linalg.norm(It[k, smp["dir_index"]] - smp["I"][j])
                    / np.linalg.norm(smp["I"][j])) for j, k in enumerate(smp["solves"])]
        out.setdefault("accuracy", {"gate": 1e-3})
        out["accuracy"]["rel_l2_vs_c_oracle_sample"] = max(eo)
        out["accuracy"]["c_oracle_sample"] = (f"{len(smp['dir_index'])} output directions, "
                                              f"solves {smp['solves']}")
        print(f"[bench] rel-L2 vs C oracle on the sample {max(eo):.2e}", file=sys.stderr, flush=True)
    solver.close()
    acc = out.get("accuracy", {})
    errs = [acc[k] for k in ("rel_l2_max", "rel_l2_vs_c_oracle_sample") if k in acc]
    rc = 0
    if not errs:
        acc["pass"] = None
        out["value"] = None
        out["value_note"] = "accuracy gate skipped (--no-direct --no-cpu): no gated value"
    else:
        acc["pass"] = bool(max(errs) <= 1e-3)
        if not acc["pass"]:
            out["value"] = None
            out["value_note"] = f"accuracy gate failed: rel-L2 {max(errs):.3e} > 1e-3"
            rc = 1
    print(json.dumps(out), flush=True)
    return rc


def bench_batch(args, W):
    """Uneven multi-segment workload through wf_batch_*: a step is one batch
    solve of every segment for all incidences (segments resident in HBM)."""
    import torch

    import wavefiber3d_amd as wf
    from wavefiber3d_amd.geometry import fiber_family_segment

    incs = W["incs"]
    nt, np_ = W["grid"]
    D = nt * np_
    S = len(incs)
    K = args.segments
    t0 = time.perf_counter()
    b = wf.Batch(n_streams=args.streams, device=0, precision=args.precision, beta=args.beta,
                 use_graph=args.graph)
    segs_n = []
    for k in range(K):
        s = fiber_family_segment(k)
        b.add(s, W["depth"])
        segs_n.append(s.n)
        if (k + 1) % 32 == 0 and k + 1 < K:
            print(f"[bench] {k + 1} segments added ({time.perf_counter() - t0:.0f} s)",
                  file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    N = float(sum(segs_n))
    print(f"[bench] {K} segments, N total {N:.3e} (min {min(segs_n)}, max {max(segs_n)}), "
          f"generate + upload + octrees {build_ms:.0f} ms", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        t0 = time.perf_counter()
        I = b.solve(incs, (nt, np_), out="torch")
        torch.cuda.synchronize()
        print(f"[bench] warm-up solve {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        I = b.solve(incs, (nt, np_), out="torch")
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / max(1, args.steps)
    info = b.info()
    pairs = N * D * S
    value = pairs / (ms_step * 1e-3)
    print(f"[bench] batch: {ms_step:.1f} ms/step, {value:.3e} pairs/s, depths "
          f"{np.bincount(info['depth']).tolist()}", file=sys.stderr, flush=True)
    out = {
        "metric": "PO point-direction pairs/s (tree vs direct) at \u22641e-3 rel-L2 vs FP64 oracle",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "bf16x3-f32acc", "data": "synthetic",
        "config": {"workload": args.workload, "description": W["desc"], "segments": K,
                   "n_points_total": N, "n_points_min": min(segs_n), "n_points_max": max(segs_n),
                   "n_directions": D, "solves_per_segment": S, "streams": args.streams,
                   "graph": args.graph, "depth_hist": np.bincount(info["depth"]).tolist()},
        "solves_per_s": K * S / (ms_step * 1e-3),
        "build_ms": build_ms,
    }
    rc = 0
    if not args.no_direct:
        # gate vs the FP64 direct sum: the largest and smallest segment on the
        # first wavelength group, or (--check-all) every segment on its first
        # solve; only the checked slices leave the GPU
        errs = {}
        lam0 = [t for t in incs if t.lambda_m == incs[0].lambda_m]
        if args.check_all:
            checks = [(k, incs[:1]) for k in range(K)]
        else:
            checks = [(k, lam0) for k in sorted({int(np.argmax(segs_n)), int(np.argmin(segs_n))})]
        t0 = time.perf_counter()
        with wf.POSolver(device=0) as sv:
            for k, chk in checks:
                s = fiber_family_segment(k)
                Id = sv.solve_fp64_direct(s, chk, (nt, np_)).reshape(len(chk), -1)
                It = I[k, :len(chk)].reshape(len(chk), -1).double().cpu().numpy()
                e = np.linalg.norm(It - Id, axis=1) / np.linalg.norm(Id, axis=1)
                errs[f"segment{k}_n{segs_n[k]}"] = float(e.max())
        emax = max(errs.values())
        out["accuracy"] = {"rel_l2_max_checked": emax, "gate": 1e-3, "pass": bool(emax <= 1e-3),
                           "checked": ("every segment, first solve" if args.check_all else
                                       "largest and smallest segment, first wavelength group")}
        if args.check_all:
            ev = np.array(list(errs.values()))
            out["accuracy"]["rel_l2_median"] = float(np.median(ev))
            out["accuracy"]["segments_checked"] = len(ev)
        else:
            out["accuracy"]["per_segment"] = errs
        if not out["accuracy"]["pass"]:
            out["value"] = None
            out["value_note"] = f"accuracy gate failed: rel-L2 {emax:.3e} > 1e-3"
            rc = 1
        print(f"[bench] accuracy: max {emax:.2e} over {len(errs)} segment checks "
              f"({(time.perf_counter() - t0):.1f} s)", file=sys.stderr, flush=True)
    b.close()
    print(json.dumps(out), flush=True)
    return rc


def bench_heightfield(args, W):
    """Geometry front-end: heightfield -> surface samples on the GPU, inputs and
    outputs resident in HBM.  One step = one wf_heightfield_points call (it
    returns after its stream finished, for the zero-vertex check); HIP events
    on the call's stream bracket each call.  Algorithmic bytes: the heightfield
    read once (8 B per vertex) + 28 B written per sample."""
    import torch

    import wavefiber3d_amd as wf

    length, spacing = 400e-6, 0.1e-6
    hx, hy, nv = wf.geometry.fiber_heightfield(40e-6, 30e-6, length, spacing, seed=0)
    numel = hx.shape[0]
    h = np.concatenate([hx.astype(np.float32).ravel(), hy.astype(np.float32).ravel()])
    unitz = length / nv
    n = numel * nv
    hd = torch.from_numpy(h).cuda()
    for _ in range(max(1, args.warmup)):
        wf.heightfield_samples(hd, numel, nv, length, unitz)
    st = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(st)
        out = wf.heightfield_samples(hd, numel, nv, length, unitz)
        b.record(st)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    dev_ms = sum(a.elapsed_time(b) for a, b in ev) / args.steps
    bytes_ = 8.0 * numel * nv + 28.0 * n
    achieved = bytes_ / (dev_ms * 1e-3) / 1e9
    res = {"metric": "heightfield samples/s (auxiliary workload)", "value": n / (wall_ms * 1e-3),
           "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall_ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": args.workload, "description": W["desc"], "numel": numel,
                      "numelvert": nv, "samples": n},
           "roofline": {"bound": "hbm", "kernel": "heightfield_points (call: flag init, kernel, "
                        "4 B readback)", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "work_per_launch": bytes_, "avg_launch_ms": dev_ms}}
    if not args.no_cpu:
        # CPU leg: the C oracle on the host (timed) and, as the checker, the
        # GPU result of the same heightfield compared bit for bit
        from oracle import oracle

        got = [t.cpu().numpy() for t in wf.heightfield_samples(hd, numel, nv, length,
                                                              float(np.float32(unitz)))]
        t0 = time.perf_counter()
        ref, _ = oracle.heightfield_points(h, numel, nv, length, np.float32(unitz))
        cs = time.perf_counter() - t0
        res["parity_bit_identical"] = bool(all(np.array_equal(a, b) for a, b in zip(got, ref)))
        res["cpu_baseline"] = {"value": n / cs, "unit": "samples/s", "cores": 1, "kind": "port",
                               "sample": f"the whole heightfield ({n} samples), C oracle, {cs:.2f} s"}
    del out
    print(f"[bench] heightfield: {wall_ms:.3f} ms/call wall, {dev_ms:.3f} ms device, "
          f"{achieved:.0f} GB/s, parity {res.get('parity_bit_identical')}", file=sys.stderr,
          flush=True)
    print(json.dumps(res), flush=True)
    return 0


def bench_average(args, W):
    """Output-grid block average on the GPU (inputs resident in HBM, device
    path, asynchronous): HIP events on the launch stream around each call.
    Algorithmic bytes: every input sample read once + every output written."""
    import torch

    import wavefiber3d_amd as wf

    g = torch.rand((25, 7200, 1440), dtype=torch.float32, device="cuda")
    for _ in range(max(1, args.warmup)):
        wf.average_grid(g, 180, 360)
    st = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(st)
        out = wf.average_grid(g, 180, 360)
        b.record(st)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    dev_ms = sum(a.elapsed_time(b) for a, b in ev) / args.steps
    bytes_ = 4.0 * g.numel() + 4.0 * out.numel()
    achieved = bytes_ / (dev_ms * 1e-3) / 1e9
    res = {"metric": "grid average input samples/s (auxiliary workload)",
           "value": g.numel() / (wall_ms * 1e-3), "unit": "samples/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall_ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic", "config": {"workload": args.workload, "description": W["desc"]},
           "roofline": {"bound": "hbm", "kernel": "average_grid_kernel", "achieved": achieved,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": None, "work_per_launch": bytes_, "avg_launch_ms": dev_ms}}
    if not args.no_cpu:
        from oracle import oracle

        got = out[0].cpu().numpy()
        x = g[0].cpu().numpy()
        t0 = time.perf_counter()
        ref = oracle.average_grid(x, 180, 360)
        cs = time.perf_counter() - t0
        res["parity_bit_identical"] = bool(np.array_equal(got, ref))
        res["cpu_baseline"] = {"value": x.size / cs, "unit": "samples/s", "cores": 1, "kind": "port",
                               "sample": f"one 7200x1440 grid, C oracle, {cs:.3f} s"}
    print(f"[bench] average: {dev_ms:.3f} ms/call device, {achieved:.0f} GB/s, parity "
          f"{res.get('parity_bit_identical')}", file=sys.stderr, flush=True)
    print(json.dumps(res), flush=True)
    return 0


def _aux_timing(call, steps):
    """Wall and HIP-event (current stream) time per call, in ms."""
    import torch

    st = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for a, b in ev:
        a.record(st)
        out = call()
        b.record(st)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    dev_ms = sum(a.elapsed_time(b) for a, b in ev) / steps
    return out, wall_ms, dev_ms


def bench_finite(args, W):
    """Finite-distance FP64 validator (wf_po_solve_fp64_finite), surface in HBM.
    One step = one solve of the reference's validator configuration.  The
    dominant kernel is the FP64 direct sum: 48 useful FP64 flops per
    point-direction pair (6 complex multiply-adds) against the FP64 peak (the
    vector and matrix FP64 peaks of MI355X are both 78.6 TFLOP/s)."""
    import torch

    import wavefiber3d_amd as wf

    g = os.path.join(ROOT, "tests", "golden")
    s = wf.read_soa_mesh(os.path.join(g, "ellipsoid_3_3_2_float.binary"),
                         os.path.join(g, "ellipsoid_3_3_2_normal_float.binary"),
                         os.path.join(g, "ellipsoid_3_3_2_area_float.binary"))
    inc = wf.Incidence(0.0, 0.0, 500e-9, "M", 1.55 + 0.3j, 1e-2)
    grid = (2, 3600)
    dev = [torch.from_numpy(a).cuda() for a in s.arrays()]
    solver = wf.POSolver(device=0)
    for _ in range(max(1, args.warmup)):
        solver.solve_fp64_finite(dev, [inc], grid, 1.0)
    I, wall_ms, dev_ms = _aux_timing(lambda: solver.solve_fp64_finite(dev, [inc], grid, 1.0),
                                     args.steps)
    k = solver.stats()["kernels"]["direct_fp64"]
    pairs = float(s.n) * grid[0] * grid[1]
    kms = k["ms"] / k["launches"]
    tflops = 48.0 * pairs / (kms * 1e-3) / 1e12
    res = {"metric": "finite-distance validator point-direction pairs/s (auxiliary workload)",
           "value": pairs / (wall_ms * 1e-3), "unit": "pairs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall_ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "reference ellipsoid mesh",
           "config": {"workload": args.workload, "description": W["desc"], "n_points": s.n,
                      "n_directions": grid[0] * grid[1]},
           "roofline": {"bound": "mfma", "kernel": "direct_fp64_kernel<6, finite> (FP64 vector ALU)",
                        "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": tflops / FP64_PEAK_TFLOPS, "traffic": None,
                        "work_per_launch": 48.0 * pairs, "avg_launch_ms": kms},
           "device_ms_per_call": dev_ms}
    if not args.no_cpu:
        from oracle import oracle

        nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        ref, _ = oracle.solve_finite(s.arrays(), [inc.as_tuple()], *grid, 1.0, nthreads=nthreads)
        cs = time.perf_counter() - t0
        got = I.cpu().numpy().reshape(-1)
        res["parity_rel_l2"] = float(np.linalg.norm(got - ref[0]) / np.linalg.norm(ref[0]))
        res["cpu_baseline"] = {"value": pairs / cs, "unit": "pairs/s", "cores": nthreads,
                               "kind": "port",
                               "sample": f"the whole workload, C oracle (OpenMP), {cs:.2f} s"}
    solver.close()
    print(f"[bench] finite: {wall_ms:.3f} ms/call wall, kernel {kms:.3f} ms, {tflops:.1f} TFLOP/s, "
          f"parity {res.get('parity_rel_l2')}", file=sys.stderr, flush=True)
    print(json.dumps(res), flush=True)
    return 0


def bench_rough2d(args, W):
    """2D cross-section PO (wf_rough2d_solve), polyline in HBM: one step = one
    call with all 50 incidences.  The dominant kernel is the far-field sum:
    48 useful FP64 flops per sample-direction pair, FP64 peak 78.6 TFLOP/s."""
    import torch

    import wavefiber3d_amd as wf

    g = os.path.join(ROOT, "tests", "golden")
    x, y = wf.read_crosssection(os.path.join(g, "ellipse_x.txt"), os.path.join(g, "ellipse_y.txt"))
    incs = [wf.Rough2dIncidence(wf.rough2d.wavelength(l), p) for p in "ME" for l in range(25)]
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    for _ in range(max(1, args.warmup)):
        wf.rough2d_solve(xd, yd, incs)
    I, wall_ms, dev_ms = _aux_timing(lambda: wf.rough2d_solve(xd, yd, incs), args.steps)
    pairs = float(x.shape[0]) * 3600 * len(incs)
    tflops = 48.0 * pairs / (dev_ms * 1e-3) / 1e12
    res = {"metric": "2D cross-section PO sample-direction pairs/s (auxiliary workload)",
           "value": pairs / (wall_ms * 1e-3), "unit": "pairs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall_ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "reference ellipse cross-section",
           "config": {"workload": args.workload, "description": W["desc"], "n_samples": x.shape[0],
                      "incidences": len(incs), "n_phi": 3600},
           "roofline": {"bound": "mfma", "kernel": "rough2d call (currents + far field, FP64 "
                        "vector ALU; HIP events around the call)", "achieved": tflops,
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": tflops / FP64_PEAK_TFLOPS, "traffic": None,
                        "work_per_launch": 48.0 * pairs, "avg_launch_ms": dev_ms}}
    if not args.no_cpu:
        from oracle import oracle

        got = I.cpu().numpy()
        t0 = time.perf_counter()
        errs = []
        for i, t in enumerate(incs):
            ref = oracle.rough2d(x, y, t.theta, t.phi, t.lambda_m, t.pol, t.eta)
            errs.append(float(np.abs(got[i] - ref).max() / np.abs(ref).max()))
        cs = time.perf_counter() - t0
        res["parity_max_rel"] = max(errs)
        res["cpu_baseline"] = {"value": pairs / cs, "unit": "pairs/s", "cores": 1, "kind": "port",
                               "sample": f"all {len(incs)} incidences, C oracle (one thread), "
                                         f"{cs:.2f} s"}
    print(f"[bench] rough2d: {wall_ms:.3f} ms/call wall, {dev_ms:.3f} ms device, "
          f"{tflops:.2f} TFLOP/s, parity {res.get('parity_max_rel')}", file=sys.stderr, flush=True)
    print(json.dumps(res), flush=True)
    return 0


def bench_adjoint(args, W):
    """Reverse mode on the 200k-point fiber: forward and adjoint times and
    gradient parity of the tree with FP64 patterns (the default gradient
    mode), the tree with FP32 patterns, and the FP64 direct sum (the
    reference gradient), at the autotuned depth.  Each gradient is checked
    against the FP64 direct-sum gradient (relative L2 of dL/d delta_n for the
    normal displacement x -> x + delta n) and against central differences of
    the FP64 forward: 64 seeded points moved along their normals by 1e-4 um
    and h/2, and the cuticle tilt along the FP64 tangent of the surface
    recipe at 0.002 and 0.001 degrees, each step exactly centred in float32
    and the ratio FD / prediction Richardson-extrapolated
    (wavefiber3d_amd/gradcheck.py).  Gate: 1e-4 relative."""
    import torch

    import wavefiber3d_amd as wf
    from wavefiber3d_amd import gradcheck as gc
    from wavefiber3d_amd.autograd import log_intensity_loss, po_intensity

    incs, grid = W["incs"], W["grid"]
    tau0 = torch.tensor(3.0, dtype=torch.float64, device="cuda")
    s64 = [t.detach() for t in gc.c5_surface(tau0)]
    s32 = [t.float().contiguous() for t in s64]
    n = s32[0].numel()
    nrm = torch.stack(s64[3:6])
    out = {"metric": "adjoint: forward + reverse time and gradient parity (auxiliary workload)",
           "unit": "ms", "n_gpus": 1, "data": "synthetic", "higher_is_better": False,
           "config": {"workload": args.workload, "description": W["desc"], "n_points": n,
                      "grid": list(grid)}}
    with wf.POSolver(device=0) as sv:
        I0 = sv.solve_fp64_direct(s32, incs, grid)
        rng = np.random.default_rng(12345)
        target = (I0 * torch.from_numpy(np.exp(0.3 * rng.standard_normal(I0.shape))).cuda()).float()

        def sync_time(fn, reps=3):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                r = fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / reps, r

        sv.set_geometry(s32, "auto")
        best = sv.autotune(incs, grid)
        ms_fwd, I = sync_time(lambda: sv.solve(incs, grid, out="torch"))
        depth = sv.stats()["depth"]
        sv.set_geometry(s32, depth)
        gI = (2.0 * (torch.log(I) - torch.log(target)) / I).contiguous()
        ms = {"tree_forward_solve": ms_fwd}
        for mode in ("tree64", "tree32"):
            ms[f"{mode}_forward"], _ = sync_time(lambda: sv.grad_forward(incs, grid, mode=mode))
            ms[f"{mode}_vjp"], _ = sync_time(lambda: sv.vjp(incs, grid, gI, mode=mode))
        ms["fp64_forward"], _ = sync_time(lambda: sv.solve_fp64_direct(s32, incs, grid))
        sv.set_geometry(s32, 0)
        ms["fp64_vjp"], _ = sync_time(lambda: sv.vjp(incs, grid, gI, mode="direct64"))
        out["ms"] = ms
        out["depth"] = {"autotuned_best": best, "used": depth}
        print(f"[bench] adjoint ms: " + ", ".join(f"{k} {v:.1f}" for k, v in ms.items()),
              file=sys.stderr, flush=True)
        grads = {m: gc.input_gradients(sv, s32, incs, grid, target, m, 0 if m == "direct64" else depth)
                 for m in ("direct64", "tree64", "tree32")}
        gd = {m: (torch.stack(g[:3]) * nrm).sum(0) for m, g in grads.items()}
        rel = {m: float(torch.linalg.norm(gd[m] - gd["direct64"]) / torch.linalg.norm(gd["direct64"]))
               for m in ("tree64", "tree32")}
        rn = gc.fd_normal_ratio(sv, s32, nrm, incs, grid, target, grads)
        rt = gc.fd_tangent_ratio(sv, gc.c5_surface, 3.0, incs, grid, target, grads)
        # dL/dtau itself, through the differentiable recipe (tree64)
        tau = torch.tensor(3.0, dtype=torch.float64, device="cuda", requires_grad=True)
        L = log_intensity_loss(po_intensity([t.float() for t in gc.c5_surface(tau)], incs, grid,
                                            depth, sv, mode="tree64"), target)
        L.backward()
    out["gradient"] = {
        "vs_fp64_gradient_rel_l2": rel,
        "vs_fd_normal_64pts_rel": {m: abs(r - 1.0) for m, r in rn.items()},
        "vs_fd_tilt_tangent_rel": {m: abs(r - 1.0) for m, r in rt.items()},
        "dL_dtau_tree64": float(tau.grad.item()),
        "gate": 1e-4,
        "pass_tree64": bool(abs(rn["tree64"] - 1) <= 1e-4 and abs(rt["tree64"] - 1) <= 1e-4),
    }
    g = out["gradient"]
    print(f"[bench] gradients vs FP64 gradient: tree64 {rel['tree64']:.2e}, tree32 {rel['tree32']:.2e}; "
          f"vs FD (normals): " + ", ".join(f"{m} {v:.2e}" for m, v in g["vs_fd_normal_64pts_rel"].items())
          + "; vs FD (tilt): " + ", ".join(f"{m} {v:.2e}" for m, v in g["vs_fd_tilt_tangent_rel"].items()),
          file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    return 0

if __name__ == "__main__":
    sys.exit(main())
