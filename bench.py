"""MPD-RDME throughput on one MI355X.

A step is one MPD-RDME timestep (x, y, z diffusion sweeps, reaction phase,
overflow resolution) over the whole lattice.  The default workload is the
whole-cell-scale lattice: 256x256x512 sites (33.5M), lambda = 10 nm, 8 byte
slots per site, 10 site types with 20% of the cytoplasm as 1-site obstacles,
100 species, 300 first/second-order reactions, 1e6 particles, tau = 4 us.
Inputs are generated (seeded, synthetic) and uploaded before the timed region;
the timed region runs only device work.

Output: one JSON line (rank 0) with site-steps/s, the roofline fraction of the
dominant kernel (HIP events on the solver's stream), and the CPU baseline:
this project's next-subvolume (NSM) port of the reference's CPU solver timed
on a bounded sample of the same model on this host.
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

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str, seed: int):
    from lattice_microbes_amd import models

    if name == "whole_cell":
        return models.whole_cell(seed=seed)
    if name == "whole_cell_reduced":
        return models.whole_cell(shape=(256, 256, 128), nS=30, R=90, n_particles=250_000, seed=seed)
    if name == "capsule":
        return models.capsule_minde(seed=seed)
    if name == "bimolecular":
        return models.bimolecular(seed=seed)
    raise SystemExit(f"unknown workload {name}")


def sweep_bytes_per_site(P: int) -> int:
    # read P particle bytes + 1 site-type byte, write P particle bytes
    return 2 * P + 1


def reaction_bytes_per_site(P: int) -> int:
    # read P particle bytes + 1 site-type byte (writes only on reacting sites)
    return P + 1


def cpu_baseline(spec, budget_s: float, seed: int):
    """NSM port on the same model, single-threaded, bounded by wall time."""
    import ctypes

    import oracle

    lib = oracle.load()
    orc = oracle.Oracle(spec, seed_top=seed)
    X, Y, Z = spec.shape
    S = np.ascontiguousarray(spec.S, dtype=np.int32).reshape(-1)
    DF = np.ascontiguousarray(spec.DF, dtype=np.float64).reshape(-1)
    P = ctypes.POINTER
    t0 = time.perf_counter()
    h = lib.orc_nsm_create(X, Y, Z, spec.P, spec.nS, spec.nST, spec.R, oracle._p(orc.types, ctypes.c_uint32),
                           oracle._p(orc.k, ctypes.c_double), oracle._p(orc.si, ctypes.c_uint32),
                           oracle._p(orc.s2i, ctypes.c_uint32), oracle._p(S, ctypes.c_int32),
                           oracle._p(DF, ctypes.c_double), oracle._p(orc.RL32, ctypes.c_uint32), spec.spacing,
                           oracle._p(orc.lattice, ctypes.c_uint32), oracle._p(orc.sites, ctypes.c_uint8),
                           seed, 1)
    if not h:
        raise RuntimeError("NSM setup failed")
    setup_s = time.perf_counter() - t0
    wall = 0.0
    chunk = 200_000
    while wall < budget_s:
        t1 = time.perf_counter()
        rc = lib.orc_nsm_run(h, 1e30, chunk, 0.0, 0, None, None)
        wall += time.perf_counter() - t1
        if rc:
            raise RuntimeError(f"NSM error {rc}")
    sim_t = lib.orc_nsm_time(h)
    events = lib.orc_nsm_events(h)
    lib.orc_nsm_destroy(h)
    eq_steps = sim_t / spec.tau
    value = spec.N * eq_steps / wall
    return {
        "value": value,
        "unit": "site-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"next-subvolume port of NextSubvolumeSolver on the same {spec.name} model "
                   f"({spec.N} subvolumes): {events} events in {wall:.1f} s wall = "
                   f"{sim_t * 1e6:.3f} us simulated = {eq_steps:.3f} tau-steps equivalent "
                   f"(setup {setup_s:.1f} s untimed)"),
        "events_per_s": events / wall,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="whole_cell")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--path", choices=("sweeps", "fused"), default="sweeps",
                    help="step implementation (results are bit-identical)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch

        n = 1 << 28  # 2^28 float32 = 1 GiB
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n, "launches": 3}), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs the MPD-RDME step on one GPU only; multi-GPU sharding is not part of "
              "this build (see the sharding design in DESIGN.md).", file=sys.stderr)
        return 2

    import lattice_microbes_amd as lmx

    if lmx.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    t0 = time.perf_counter()
    spec = build_workload(args.workload, args.seed)
    log(f"workload {spec.name} {spec.shape} N={spec.N} nS={spec.nS} R={spec.R} built in {time.perf_counter() - t0:.1f}s")
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, replicate=0, seed=args.seed)
    if args.path == "fused":
        solver.setPath(lmx._lib.PATH_FUSED)
    log("model uploaded")
    if args.warmup:
        solver.runTimesteps(args.warmup)
    log(f"warmup {args.warmup} steps done")
    total_ms, per = solver.timeSteps(args.steps, per_kernel=True)
    log(f"timed {args.steps} steps: {total_ms:.2f} ms")
    counts = solver.getSpeciesCounts()
    ovf_entries, ovf_steps = solver.getOverflowStats()

    N, P, K = spec.N, spec.P, args.steps
    value = N * K / (total_ms / 1e3)
    # per-kernel averages and algorithmic bytes per launch
    kernels = {}
    for name, (ms, launches) in per.items():
        if name.startswith("sweep") or name == "fused":
            b = sweep_bytes_per_site(P) * N
        elif name == "reaction":
            b = reaction_bytes_per_site(P) * N
        else:
            b = 0
        avg_ms = ms / max(launches, 1)
        kernels[name] = {"avg_us": avg_ms * 1e3, "launches": launches,
                         "alg_bytes": b, "GBps": (b / (avg_ms / 1e3) / 1e9) if b and avg_ms > 0 else None}
    dom = max((k for k in kernels if kernels[k]["alg_bytes"]), key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    dk = kernels[dom]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(f"{spec.name}_{spec.shape[0]}x{spec.shape[1]}x{spec.shape[2]}:{dom}")
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": dk["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": dk["GBps"] / HBM_PEAK_GBS, "traffic": traffic,
                "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_us": dk["avg_us"]}
    # Unfused-equivalent bytes: three sweeps plus the reaction read.
    step_alg = (3 * sweep_bytes_per_site(P) + (reaction_bytes_per_site(P) if spec.R else 0)) * N
    step_ms = total_ms / K
    out = {
        "metric": "site-steps/s & % HBM roofline; ensemble sims/hr, p50/p99 latency vs plain",
        "value": value,
        "unit": "site-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"{spec.name} {spec.shape[0]}x{spec.shape[1]}x{spec.shape[2]}",
                   "sites": N, "slots_per_site": P, "species": spec.nS, "reactions": spec.R,
                   "site_types": spec.nST, "particles": int(spec.initial_counts().sum()),
                   "lambda_m": spec.spacing, "tau_s": spec.tau, "path": args.path},
        "roofline": roofline,
        "step_roofline": {"alg_bytes_per_step": step_alg, "achieved_GBps": step_alg / (step_ms / 1e3) / 1e9,
                          "frac": step_alg / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS},
        "kernels": kernels,
        "overflow": {"entries": ovf_entries, "steps_with_overflow": ovf_steps},
        "final_particles": int(counts.sum()),
    }
    solver.close()
    if not args.no_cpu:
        log(f"cpu baseline ({args.cpu_budget:.0f}s budget)")
        out["cpu_baseline"] = cpu_baseline(spec, args.cpu_budget, args.seed)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
