#!/usr/bin/env python3
"""Benchmark: plain FP64 CG (TeaLeaf) at 4096^2 on one MI355X.

Workload (one "step" = one TeaLeaf timestep, i.e. one full CG solve): the
4096 x 4096 deck with K = 1/density, the tea.in rectangles plus 16 seeded
random rectangles, eps 1e-15 absolute, max_iters 10000, dt 0.004.  W warm-up
timesteps run first; the initial state is then restored and K timesteps are
timed (device-synchronised on both sides) by bin/tealeaf itself.

A second part times the solve the metric names: one seeded 4096 x 4096
field of 64 rectangles (log10 density uniform in [-2, 2], K = 1/density,
seed 20269202, not among the training problems), one timestep solved
to 1e-10 relative residual by plain FP64 CG, mixed-precision CG, Jacobi-PCG,
flexible CG with the two-level preconditioner and a non-learned coarse solve
(6 Jacobi sweeps, the network's training target; 1 sweep) and flexible CG
with the learned preconditioner; each after one warm-up step, with its
iterations, time and final-field error against FP64 CG.  The same field
drawn at 8192 x 8192 is then solved by plain FP64 CG and learned FCG.

Printed (one JSON line): effective HBM throughput of the timed solve,
  value = 88 B x N^2 x (CG iterations executed) / t_solve,
88 B being the algorithmic FP64 traffic per cell per iteration (11 words),
plus time-to-solution, iterations/s, cell-updates/s, the roofline of the
dominant kernel (cg_fused_w: 8 words = 64 B per cell per launch, timed with
HIP events on the library's stream; also set against a measured 16-byte
device-copy rate) and a CPU baseline (the OpenMP CPU restatement of the
reference kernels, unfused 13-word dataflow, timed on a bounded sample of the
same problem on this host).
"""

import argparse
import json
import os
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("time-to-solution @4096², 1e-10 rel. residual: learned FCG vs plain CG; "
          "HBM GB/s")
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_ITER = 88      # fused FP64 iteration, algorithmic
BYTES_FUSED_W = 64            # cg_fused_w: p_old, r, u, kx, ky read; p, u, w written
BYTES_FUSED_R = 24            # cg_fused_r: r, w read; r written
PROFILE_TRAFFIC = os.path.join(REPO, "profiles", "cg_fused_w_4096_traffic.json")


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed timesteps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed timesteps")
    ap.add_argument("--cells", type=int, default=4096, help="grid is cells x cells")
    ap.add_argument("--max-iters", type=int, default=10000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU-baseline sample length (0 disables it)")
    ap.add_argument("--unfused", action="store_true",
                    help="one launch per reference kernel (13 words/iteration)")
    ap.add_argument("--tts", type=int, default=1,
                    help="run the time-to-solution comparison (0 skips it)")
    ap.add_argument("--tts-cells", type=int, default=0,
                    help="grid of the time-to-solution comparison (default: --cells)")
    ap.add_argument("--tts-max-iters", type=int, default=12000,
                    help="iteration cap of the time-to-solution solves")
    ap.add_argument("--tts-large", type=int, default=8192,
                    help="grid of the second time-to-solution comparison, plain FP64 CG "
                         "against learned FCG (0 skips it)")
    ap.add_argument("--solvers", type=int, default=1,
                    help="time CG, Chebyshev, PPCG and Jacobi on the known-answer "
                         "deck (0 skips it)")
    ap.add_argument("--solver-max-iters", type=int, default=0,
                    help="cap on every solve of that comparison (0: none; for profiling)")
    return ap.parse_args()


def copy_peak(gib=1.0, reps=20):
    """Measured streaming ceiling: a 16-byte nontemporal device copy of
    `gib` GiB (libteahip's tea_hip_copy_bandwidth), read + write bytes per
    second, to set beside the 8 TB/s peak."""
    import ctypes
    from exploringsycl_amd import native
    lib = native.load()
    gbs = ctypes.c_double()
    native.check(lib.tea_hip_copy_bandwidth(0, int(gib * (1 << 30)), reps,
                                            ctypes.byref(gbs)), "copy_bandwidth")
    return gbs.value


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _omp_baseline_lib():
    """The c_kernels-style OpenMP CG (oracle/ckernels_omp.c) built for this
    host's vector ISA: the AVX-512 build where /proc/cpuinfo lists avx512f
    (what -march=native selects there), else the AVX2 build."""
    import ctypes
    flags = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = line
                    break
    except OSError:
        pass
    isa = "v4" if " avx512f" in flags else "v3"
    lib = ctypes.CDLL(os.path.join(REPO, "oracle", "build", f"libckernels_omp_{isa}.so"))
    d, i, l = ctypes.c_double, ctypes.c_int, ctypes.c_long
    pd, pi = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    lib.omp_cg_run.argtypes = [i, i, i, i, d, d, d, i, i, l, d, pd, pd, pi, pi, pd]
    lib.omp_cg_run.restype = d
    lib.omp_threads.restype = i
    return lib, isa


def _omp_run(lib, deck, coefficient, max_total=0):
    """Runs deck's timesteps (at most max_total CG iterations in all) on the
    OpenMP baseline from the deck's initial state (set up by the oracle's
    state kernels); returns (seconds, per-step executed iterations, temp)."""
    import ctypes
    import numpy as np
    from tests.oracle_ctypes import Oracle, OracleSolver
    o = Oracle()
    s = OracleSolver(o, deck.x_cells, deck.y_cells, deck.states, coefficient=coefficient,
                     dt=deck.dt, extent=deck.extent)
    dens = np.ascontiguousarray(s.f["density"])
    ener = np.ascontiguousarray(s.f["energy"])
    iters = np.zeros(deck.end_step, dtype=np.int32)
    steps = ctypes.c_int(0)
    temp = ctypes.c_double(0.0)
    pd = ctypes.POINTER(ctypes.c_double)
    secs = lib.omp_cg_run(s.x, s.y, s.h, coefficient, s.rx, s.ry, deck.eps, deck.max_iters,
                          deck.end_step, max_total, s.dx * s.dy, dens.ctypes.data_as(pd),
                          ener.ctypes.data_as(pd),
                          iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                          ctypes.byref(steps), ctypes.byref(temp))
    return secs, [int(v) for v in iters[:steps.value]], temp.value


def cpu_baseline(cells, seconds):
    """The reference's CPU path in the style of its OpenMP c_kernels
    (oracle/ckernels_omp.c: unfused calc_w / calc_ur / calc_p, OpenMP
    reductions over rows, -O3 -march=native -funroll-loops -fopenmp) on all
    the host threads OpenMP is given: a bounded sample of the same 4096^2
    deck (as many CG iterations of its first timestep as fit in about
    `seconds`), and the CPU configuration (512^2, 10 timesteps, K =
    density) in full."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from exploringsycl_amd.deck import config_deck
    lib, isa = _omp_baseline_lib()
    threads = lib.omp_threads()
    deck = config_deck(2, x_cells=cells, end_step=1)
    # Calibrate on 20 iterations, then size the sample to `seconds`.
    t, it, _ = _omp_run(lib, deck, 2, max_total=20)
    rate = sum(it) / t
    n = max(20, int(rate * seconds))
    t, it, _ = _omp_run(lib, deck, 2, max_total=n)
    iters = sum(it)
    gbs = BYTES_PER_CELL_ITER * cells * cells * iters / t / 1e9
    c1 = config_deck(1)
    t1, it1, temp1 = _omp_run(lib, c1, 1)
    return {
        "value": round(gbs, 3), "unit": "GB/s",
        "cores": threads, "kind": "port",
        "sample": (f"first {iters} CG iterations of timestep 1 of the same {cells}^2 deck, "
                   f"{t:.1f} s, OpenMP c_kernels-style CG (unfused, FP64, "
                   f"-march=x86-64-{isa}), same 88 B/cell/iteration convention; "
                   f"{iters / t:.2f} iterations/s"),
        "iterations_per_s": iters / t,
        "config1_512x512x10": {
            "seconds": round(t1, 4), "iterations": it1, "temp": temp1,
            "cg_iterations_per_s": round(sum(it1) / t1, 2),
            "gbs": round(BYTES_PER_CELL_ITER * 512 * 512 * sum(it1) / t1 / 1e9, 3)},
    }


ALL_TTS_SOLVERS = ("plain_fp64_cg", "mixed_precision_cg", "jacobi_pcg",
                   "two_level_jacobi_sweeps_6", "two_level_jacobi_sweeps_1",
                   "learned_fcg")


def time_to_solution(n, max_iters=12000, timeout=1200, names=ALL_TTS_SOLVERS):
    """Plain FP64 CG, MPCG, Jacobi-PCG, the two-level flexible CG with a
    non-learned coarse solve (6 Jacobi sweeps, the network's training
    target, and 1 sweep, y = D^-1 R r) and learned FCG on one unseen field
    of the 64-rectangle workload at n x n, stopped at 1e-10 relative
    residual (the solvers in `names`)."""
    import numpy as np
    import shutil
    from exploringsycl_amd.deck import Deck, problem_seed
    from exploringsycl_amd.learned import WEIGHTS_PATH
    from exploringsycl_amd.native import TEALEAF_BIN
    from exploringsycl_amd.runner import run_deck

    seed = problem_seed(5, 3000)
    extras = {"plain_fp64_cg": [], "mixed_precision_cg": ["use_mpcg"],
              "jacobi_pcg": ["use_jacobi_pcg"],
              "two_level_jacobi_sweeps_6": ["use_fcg", "fcg_coarse_sweeps=6"],
              "two_level_jacobi_sweeps_1": ["use_fcg", "fcg_coarse_sweeps=1"],
              "learned_fcg": ["use_fcg", "fcg_weights=net.bin"]}
    solvers = [(k, extras[k]) for k in names
               if k != "learned_fcg" or os.path.exists(WEIGHTS_PATH)]
    out = {"workload": (f"{n}x{n} cells, 64 seeded rectangles (log10 density "
                        f"U[-2,2], seed {seed}) + tea.in states, K=1/density, "
                        f"1 timestep, rel. residual 1e-10 (max {max_iters} iterations), "
                        "after 1 warm-up step"),
           "solvers": {}}
    fields = {}
    with tempfile.TemporaryDirectory(prefix="tea_tts_") as root:
        for name, extra in solvers:
            deck = Deck.with_random_rectangles(
                n, n, 64, seed, density="log10", density_range=(-2.0, 2.0),
                end_step=1, coefficient="inverse", max_iters=max_iters,
                extra=["rel_eps=1e-10"] + extra)
            wd = os.path.join(root, name)
            os.makedirs(wd)
            if name == "learned_fcg":
                shutil.copy(WEIGHTS_PATH, os.path.join(wd, "net.bin"))
            rec, _ = run_deck(deck, binary=TEALEAF_BIN, workdir=wd, timeout=timeout,
                              extra_args=["--warmup-steps", "1", "--field-out", "u.bin"])
            st = rec["steps"][0]
            progress(f"{n}^2 {name}: {st['iterations']} iterations, "
                     f"{rec['timed_seconds']:.3f} s")
            fields[name] = np.fromfile(os.path.join(wd, "u.bin"))
            out["solvers"][name] = {
                "seconds": round(rec["timed_seconds"], 4),
                "iterations": st["iterations"],
                "rel_residual": st["rel_residual"],
                "fallback_at": st["fallback_at"],
                "temp": rec["summaries"][-1]["temp"],
            }
    ref = fields["plain_fp64_cg"]
    tref = out["solvers"]["plain_fp64_cg"]
    for name, rec in out["solvers"].items():
        d = fields[name] - ref
        rec["field_err_max_rel"] = float(np.abs(d).max() / np.abs(ref).max())
        rec["field_err_l2_rel"] = float(np.linalg.norm(d) / np.linalg.norm(ref))
        rec["temp_rel_err"] = abs(rec["temp"] / tref["temp"] - 1.0)
        rec["speedup_vs_plain_cg"] = round(tref["seconds"] / rec["seconds"], 3)
    return out


# Algorithmic bytes per interior cell of each solver's streaming pass.
SOLVER_KERNELS = {"cg": ("cg_fused_w", BYTES_FUSED_W),
                  "chebyshev": ("cheby_two_step", 72),  # two steps per launch
                  "ppcg": ("ppcg_two_step", 64),   # two inner steps per launch
                  "jacobi": ("jacobi_fused", 40)}


def other_solvers(n, jacobi_iters=2000, max_iters=0, timeout=900):
    """The reference's solvers on its own known-answer problem (tea.problems
    row n n 1: tea.in states, K = density, eps 1e-15 on the solver's own
    test): time to solution after one warm-up step, iteration counts, the
    known-answer check, and the average duration and bandwidth of each
    solver's streaming pass (HIP events around that kernel only).  Jacobi
    does not converge within max_iters at this size; it is capped and timed
    per sweep.  max_iters > 0 caps every solve (for profiling runs)."""
    from exploringsycl_amd.deck import Deck
    from exploringsycl_amd.native import TEALEAF_BIN
    from exploringsycl_amd.runner import run_deck

    out = {"workload": (f"{n}x{n} cells, tea.in states, coefficient_density, 1 "
                        "timestep (the tea.problems known-answer deck), eps 1e-15, "
                        f"Jacobi capped at {jacobi_iters} sweeps, after 1 warm-up step"),
           "solvers": {}}
    cap = [f"max_iters {max_iters}"] if max_iters > 0 else []
    extras = {"cg": cap, "chebyshev": ["use_chebyshev"] + cap, "ppcg": ["use_ppcg"] + cap,
              "jacobi": ["use_jacobi", f"max_iters {max_iters or jacobi_iters}"]}
    for name, extra in extras.items():
        kname, kbytes = SOLVER_KERNELS[name]
        with tempfile.TemporaryDirectory(prefix="tea_solver_") as wd:
            stats_path = os.path.join(wd, "kernel_stats.json")
            rec, _ = run_deck(Deck(n, n, end_step=1, extra=extra), binary=TEALEAF_BIN,
                              workdir=wd, timeout=timeout,
                              env={"TEA_HIP_STATS": stats_path,
                                   "TEA_HIP_STATS_ONLY": kname},
                              extra_args=["--warmup-steps", "1"])
            with open(stats_path) as fp:
                k = json.load(fp)[kname]
        st = rec["steps"][-1]
        avg_ms = k["total_ms"] / k["launches"]
        gbs = kbytes * n * n / (avg_ms * 1e-3) / 1e9
        out["solvers"][name] = {
            "seconds": round(rec["timed_seconds"], 4),
            "iterations": st["iterations"],
            "after_switch": st["solver_iterations"],
            "known_answer_passed": rec["check_passed"] if rec["check_found"] else None,
            "temp": rec["summaries"][-1]["temp"],
            "kernel": {"name": kname, "launches": k["launches"],
                       "avg_us": round(avg_ms * 1e3, 2),
                       "algorithmic_bytes_per_cell": kbytes,
                       "achieved_gbs": round(gbs, 1),
                       "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)},
        }
    return out


def progress(msg):
    """One line on stderr per finished solve (stdout carries only the JSON line)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def main():
    args = parse_args()
    if args.gpus != 1:
        print("bench.py: this build drives one GPU; multi-GPU runs are not part "
              "of it (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from exploringsycl_amd.deck import config_deck
    from exploringsycl_amd.native import TEALEAF_BIN
    from exploringsycl_amd.runner import run_deck

    n = args.cells
    deck = config_deck(2, x_cells=n, end_step=args.steps)
    deck.max_iters = args.max_iters
    kname, kbytes = (("cg_calc_w", 32) if args.unfused else ("cg_fused_w", BYTES_FUSED_W))
    with tempfile.TemporaryDirectory(prefix="tea_bench_") as wd:
        stats_path = os.path.join(wd, "kernel_stats.json")
        # HIP events only around the roofline kernel, so that timing adds as
        # little as possible to the whole-solve clock.
        env = {"TEA_HIP_STATS": stats_path, "TEA_HIP_STATS_ONLY": kname,
               "TEA_HIP_FUSED": "0" if args.unfused else "1"}
        t_wall = time.perf_counter()
        rec, _ = run_deck(deck, binary=TEALEAF_BIN, workdir=wd, env=env,
                          extra_args=["--warmup-steps", str(args.warmup)],
                          timeout=3000)
        t_wall = time.perf_counter() - t_wall
        with open(stats_path) as fp:
            stats = json.load(fp)

    iters = sum(s["executed"] for s in rec["steps"])
    t = rec["timed_seconds"]
    value = BYTES_PER_CELL_ITER * n * n * iters / t / 1e9
    progress(f"roofline run: {iters} CG iterations in {t:.3f} s, {value:.1f} GB/s")

    k = stats[kname]
    avg_ms = k["total_ms"] / k["launches"]
    achieved = kbytes * n * n / (avg_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC) and not args.unfused and n == 4096:
        with open(PROFILE_TRAFFIC) as fp:
            traffic = json.load(fp).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": kbytes * n * n,
                "avg_launch_ms": round(avg_ms, 5)}

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded splitmix64 rectangles, no dataset)",
        "config": {
            "workload": (f"plain FP64 CG, {n}x{n} cells, {args.steps} timesteps, "
                         "K=1/density, tea.in + 16 seeded rectangles, eps 1e-15 "
                         f"absolute, max_iters {args.max_iters}"),
            "cells": n, "timesteps": args.steps, "seed": 20263202,
            "path": "unfused (13 words/iter)" if args.unfused else "fused (11 words/iter)",
        },
        "time_to_solution_s": round(t, 4),
        "cg_iterations": iters,
        "cg_iterations_per_s": round(iters / t, 2),
        "cell_updates_per_s": n * n * iters / t,
        "us_per_iteration": round(1e6 * t / max(iters, 1), 2),
        "iterations_per_step": [s["executed"] for s in rec["steps"]],
        "final_temp": rec["summaries"][-1]["temp"],
        "kernels_ms": {kk: round(v["total_ms"], 3) for kk, v in stats.items()},
        "roofline": roofline,
        "wall_s_including_setup": round(t_wall, 2),
    }
    if args.tts:
        tts = time_to_solution(args.tts_cells or n, args.tts_max_iters)
        line["time_to_solution"] = tts
        sv = tts["solvers"]
        line["learned_fcg_vs_plain_cg_speedup"] = (
            sv["learned_fcg"]["speedup_vs_plain_cg"] if "learned_fcg" in sv else None)
    if args.tts and args.tts_large:
        big = time_to_solution(args.tts_large, max_iters=8 * args.tts_max_iters,
                               names=("plain_fp64_cg", "learned_fcg"))
        line["time_to_solution_large"] = big
        sv = big["solvers"]
        line["learned_fcg_vs_plain_cg_speedup_large"] = (
            sv["learned_fcg"]["speedup_vs_plain_cg"] if "learned_fcg" in sv else None)
    if args.solvers:
        line["other_solvers"] = other_solvers(n, max_iters=args.solver_max_iters)
        progress("the reference's other solvers done")
    if args.cpu_seconds > 0:
        line["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
        line["cpu_baseline"]["cpu_model"] = cpu_model()
        progress("CPU baseline done")
    peak = copy_peak()
    roofline["copy_peak"] = {"value": round(peak, 1), "unit": "GB/s",
                             "method": "16-byte nontemporal device copy of 1 GiB, "
                                       "read + write bytes / time"}
    roofline["frac_of_copy_peak"] = round(roofline["achieved"] / peak, 4)
    for rec in line.get("other_solvers", {}).get("solvers", {}).values():
        rec["kernel"]["frac_of_copy_peak"] = round(rec["kernel"]["achieved_gbs"] / peak, 4)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
