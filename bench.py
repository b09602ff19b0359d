"""Benchmark: FP64 cell-updates/s of the 2000x2000 four-quadrant solve on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--extras]

Workload: the four-quadrant Riemann problem (test2D.cu:46-80) on a
2000 x 2000 FP64 mesh (two ghost layers included), [0,1]^2, gamma 1.4,
CFL 0.5, alpha 1.9, outflow ghosts, to t_end = 0.3 (about 3,000 time steps).
A "step" is one full time step of the split scheme over the whole mesh.
The state is uploaded before the timed region; the timed region is the
device-resident time loop bracketed by HIP events on the solver's stream
(the region of src/wrappers.cuh:193-288).  --steps K > 0 caps the timed loop
at K steps; the default (0) runs the whole solve to t_end.

Printed: one JSON line with the headline metric, the roofline record of the
step kernel and the CPU baseline (the FP64 host oracle on a bounded sample).
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
CU_COUNT = 256               # MI355X compute units (4 SIMDs each)
# FP64 vector peak 78.6 TFLOP/s = 256 CUs x 4 SIMDs x 16 lanes x 2 flops x 2.4 GHz:
# one wave64 v_fma_f64 per SIMD every 4 clocks
FP64_ISSUE_PEAK = 6.0e8      # wave-level FP64 VALU instructions per SIMD per second
BYTES_PER_CELL_UPDATE = 288  # 4 sweeps x (32 B read + 32 B write) + 32 B partner read
COPY_PROBE_BYTES = 1 << 28   # 256 MiB read + 256 MiB written per copy launch


def cpu_baseline(U0, steps: int, threads: int):
    """The oracle (FP64 OpenMP host restatement) on the first ``steps`` steps
    of the same workload."""
    import oracle
    _, n, _, ms = oracle.solve(U0, nthreads=threads, t_end=0.3, max_steps=steps)
    cells = U0.shape[0] * U0.shape[1]
    return cells * n / (ms / 1000.0), n, ms


def counter_record():
    """Per-launch hardware counters of the step kernel at 2000^2 (HBM bytes
    from FETCH_SIZE / WRITE_SIZE scaled by the copy-probe calibration, VALU
    instructions from SQ_INSTS_VALU), committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "step_kernel_2000_traffic.json")
    if not os.path.exists(path):
        return {}
    try:
        return json.load(open(path))
    except Exception:
        return {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (0 = full solve to t_end)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2000)
    ap.add_argument("--mode", choices=["fast", "exact"], default="fast")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--launch", default="",
                    help="strip_rows,waves_per_block,persistent: fixed geometry instead of autotuning")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the drop-in, batched, adjoint, batched-adjoint, 8192^2 and published-run legs")
    ap.add_argument("--leg", choices=["dropin", "batched", "adjoint", "batched_adjoint",
                                      "batched_adjoint_riemann", "peak_8192", "published",
                                      "published_record"],
                    help="run only this extra leg and print its record")
    ap.add_argument("--adj-fd-cells", type=int, default=50,
                    help="cells of the adjoint leg's central-difference check (0 skips it)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only run the VALU throughput probe (instructions per SIMD per second)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the plain 256 MiB device copy that calibrates the HBM byte counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: fvm2d solves on one GPU only (asked for --gpus {args.gpus}); "
              "a row-strip decomposition over several GPUs is designed but not implemented",
              file=sys.stderr)
        sys.exit(2)

    fvm = importlib.import_module("2d_fvm_cuda_amd")
    problems = fvm.problems
    if not fvm.device_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)

    if args.valu_probe:
        names = ["v_fma_f64", "v_mul_f64", "v_add_f64", "v_rcp_f64", "int32"]
        print(json.dumps({"valu_per_simd_per_s": {nm: fvm.valu_probe(k) for k, nm in enumerate(names)}}))
        return

    if args.copy_probe:
        n_doubles = COPY_PROBE_BYTES // 8
        ms = fvm.hbm_copy_probe(n_doubles, 5)
        print(json.dumps({"copy_probe_bytes_read": COPY_PROBE_BYTES, "copy_probe_bytes_written": COPY_PROBE_BYTES,
                          "avg_ms": ms, "GB_s": 2 * COPY_PROBE_BYTES / (ms / 1000.0) / 1e9}))
        return

    if args.leg:
        extras = importlib.import_module("2d_fvm_cuda_amd.benchmarks")
        fn = {"dropin": lambda f: extras.dropin(f, args.n, args.steps or -1, fvm.MODE_FAST),
              "batched": extras.batched,
              "adjoint": lambda f: extras.adjoint(f, fd_cells=args.adj_fd_cells),
              "batched_adjoint": extras.batched_adjoint,
              "batched_adjoint_riemann": lambda f: extras.batched_adjoint(f, recipe="riemann"),
              "peak_8192": extras.peak, "published": extras.published,
              "published_record": extras.published_record}[args.leg]
        print(json.dumps({args.leg: fn(fvm)}))
        return

    N = args.n
    mode = fvm.MODE_EXACT if args.mode == "exact" else fvm.MODE_FAST
    U0 = problems.initialize_4shocks(N)
    prm = dict(t_end=0.3, cfl=0.5, alpha=1.9, gamma=1.4, strategy=0, mode=mode,
               max_steps=(args.steps if args.steps > 0 else -1))
    s = fvm.Solver(N, N, **prm)
    s.upload(U0)
    tuned_ms = None
    if args.launch:
        r, w, p = (int(x) for x in args.launch.split(","))
        s.set_launch(r, w, p)
    elif not args.no_autotune:
        tuned_ms = s.autotune(steps_per_trial=4)
    launch = s.get_launch()

    # warm-up: W steps from the initial data, then restore it
    if args.warmup > 0:
        s.begin()
        s.advance(args.warmup)
        s.sync()
        s.upload(U0)

    steps, t_final, ms = s.run()
    s.close()
    cells = N * N
    value = cells * steps / (ms / 1000.0)
    # The timed region launches the step kernel once per step (graph replays
    # on the solver's stream, bracketed by HIP events), so its mean launch
    # duration is ms / steps.  Beside it: back-to-back launches on a fresh
    # copy of the workload with no step limit (every launch does a full step).
    k_avg_ms = ms / max(steps, 1)
    kt = fvm.Solver(N, N, **dict(prm, t_end=1e30, max_steps=-1))
    kt.set_launch(*launch)
    kt.upload(U0)
    kt.begin()
    kt.advance(5)
    kt.sync()
    k_b2b_ms, k_min_ms = kt.time_kernel(50)
    kt.close()

    achieved = BYTES_PER_CELL_UPDATE * cells / (k_avg_ms / 1000.0) / 1e9
    ctr = counter_record() if (N == 2000 and args.mode == "fast") else {}
    traffic = ctr.get("bytes_per_launch")
    valu_insts = ctr.get("valu_insts_per_launch")
    roof = {"bound": "fp64-valu", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "fvm_step_kernel", "kernel_avg_ms": k_avg_ms,
            "kernel_back_to_back_ms": k_b2b_ms, "kernel_min_ms": k_min_ms,
            "bytes_per_cell_update": BYTES_PER_CELL_UPDATE}
    if traffic:
        gbs = traffic / (k_avg_ms / 1000.0) / 1e9
        roof["hbm_measured"] = {"bytes_per_launch": traffic, "GB_s": gbs, "frac": gbs / HBM_PEAK_GBS}
    if valu_insts:
        per_simd = valu_insts / (k_avg_ms / 1000.0) / (4 * CU_COUNT)
        roof["valu"] = {"insts_per_launch": valu_insts, "per_simd_per_s": per_simd,
                        "fp64_issue_peak_per_simd_per_s": FP64_ISSUE_PEAK,
                        "issue_frac": per_simd / FP64_ISSUE_PEAK}
    out = {
        "metric": "FP64 cell-updates/s at 2000^2 (% HBM roofline); batched and adjoint vs plain",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms / max(steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"four-quadrant Riemann problem {N}x{N} FP64 to t_end 0.3 "
                               "(CFL 0.5, alpha 1.9, gamma 1.4, outflow)",
                   "mesh": [N, N], "t_end": 0.3, "t_final": t_final, "mode": args.mode,
                   "launch": {"strip_rows": launch[0], "waves_per_block": launch[1],
                              "persistent": launch[2]}},
        "roofline": roof,
    }
    if tuned_ms is not None:
        out["config"]["autotune_best_ms_per_step"] = tuned_ms

    if not args.no_extras:
        extras = importlib.import_module("2d_fvm_cuda_amd.benchmarks")
        legs = (("dropin", lambda f: extras.dropin(f, N, steps, mode, out["ms_per_step"])),
                ("batched", extras.batched),
                ("adjoint", lambda f: extras.adjoint(f, fd_cells=args.adj_fd_cells)),
                ("batched_adjoint", extras.batched_adjoint),
                ("batched_adjoint_riemann", lambda f: extras.batched_adjoint(f, recipe="riemann")),
                ("peak_8192", extras.peak), ("published", extras.published))
        for key, fn in legs:
            try:
                out[key] = fn(fvm)
            except Exception as e:  # a failing leg is reported, not fatal to the headline
                out[key] = {"error": f"{type(e).__name__}: {e}"}
            print(f"bench: {key} done", file=sys.stderr, flush=True)

    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
        t0 = time.time()
        v, n, cms = cpu_baseline(U0, args.cpu_steps, threads)
        out["cpu_baseline"] = {"value": v, "unit": "cell-updates/s", "cores": threads,
                               "kind": "port",
                               "sample": f"oracle, first {n} steps of the same {N}x{N} workload "
                                         f"({cms / 1000.0:.1f} s wall)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
