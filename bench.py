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
import ctypes
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


def list_entry_bytes(P: int) -> int:
    # one reaction-list entry: site index, P particle bytes, site-type byte
    return 4 + P + 1


def uses_reaction_list(spec) -> bool:
    """The default step lists the z sweep's occupied output sites and runs the
    reaction phase on them only: the 4-sites-per-thread kernels (X % 4 == 0),
    reactions present, none of them zero-order."""
    from lattice_microbes_amd.models import ZEROTH

    return spec.shape[0] % 4 == 0 and spec.R > 0 and not bool((spec.types == ZEROTH).any())


def kernel_source_digest() -> str:
    """sha256 over the device sources (csrc/*.hip and csrc/lmx_*.h, sorted by
    name): the kernels the HBM counters in profiles/traffic.json were
    collected from."""
    import glob
    import hashlib

    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "lattice_microbes_amd", "csrc")
    for path in sorted(glob.glob(os.path.join(csrc, "*.hip")) + glob.glob(os.path.join(csrc, "lmx_*.h"))):
        h.update(os.path.basename(path).encode())
        h.update(open(path, "rb").read())
    return h.hexdigest()


def measured_traffic(key: str):
    """HBM bytes per launch of one kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    scaled by a plain copy's counters) from profiles/traffic.json, used only
    if they were collected from the current kernel sources."""
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tfile):
        return None, "no counters committed"
    try:
        data = json.load(open(tfile))
    except (OSError, ValueError):
        return None, "profiles/traffic.json unreadable"
    if data.get("kernel_sources_sha256") != kernel_source_digest():
        return None, "counters were collected from other kernel sources: dropped"
    value = data.get("bytes_per_launch", {}).get(key)
    if value is None:
        return None, f"no counters for {key}"
    return value, f"rocprofv3 FETCH_SIZE/WRITE_SIZE at kernel sources {data['kernel_sources_sha256'][:16]}"


def host_cpu() -> dict:
    """The host's CPU model and core counts (os.cpu_count and this process's
    affinity)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "affinity": avail}


def nsm_run(spec, seed: int, replicate: int, budget_s: float):
    """The next-subvolume port (NextSubvolumeSolver, one thread) on one
    replicate of spec for about budget_s of wall time: (simulated seconds,
    events, wall seconds, setup seconds).  The C port runs without the GIL, so
    several calls proceed in parallel from Python threads."""
    import ctypes

    import oracle

    lib = oracle.load()
    orc = oracle.Oracle(spec, seed_top=seed, replicate=replicate)
    X, Y, Z = spec.shape
    S = np.ascontiguousarray(spec.S, dtype=np.int32).reshape(-1)
    DF = np.ascontiguousarray(spec.DF, dtype=np.float64).reshape(-1)
    t0 = time.perf_counter()
    h = lib.orc_nsm_create(X, Y, Z, spec.P, spec.nS, spec.nST, spec.R, oracle._p(orc.types, ctypes.c_uint32),
                           oracle._p(orc.k, ctypes.c_double), oracle._p(orc.si, ctypes.c_uint32),
                           oracle._p(orc.s2i, ctypes.c_uint32), oracle._p(S, ctypes.c_int32),
                           oracle._p(DF, ctypes.c_double), oracle._p(orc.RL32, ctypes.c_uint32), spec.spacing,
                           oracle._p(orc.lattice, ctypes.c_uint32), oracle._p(orc.sites, ctypes.c_uint8),
                           seed, replicate + 1)
    if not h:
        raise RuntimeError("NSM setup failed")
    setup_s = time.perf_counter() - t0
    wall = 0.0
    chunk = 200_000
    try:
        while wall < budget_s:
            t1 = time.perf_counter()
            rc = lib.orc_nsm_run(h, 1e30, chunk, 0.0, 0, None, None)
            wall += time.perf_counter() - t1
            if rc:
                raise RuntimeError(f"NSM error {rc}")
        return lib.orc_nsm_time(h), lib.orc_nsm_events(h), wall, setup_s
    finally:
        lib.orc_nsm_destroy(h)


# Kernels of the default whole-cell step, by rocprofv3 kernel name prefix.
TRAFFIC_KERNELS = {"sweep_x": "void lmx::sweep_x4_kernel<2>", "sweep_y": "void lmx::sweep_tile_kernel<2, false>",
                   "sweep_z": "void lmx::sweep_tile_kernel<2, true>", "reaction": "void lmx::reaction_list_kernel<2>"}


def counter_means(path: str, counter: str) -> dict:
    """Mean of one counter per dispatch, by kernel name, from a rocprofv3
    counter_collection.csv."""
    import csv

    acc = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            acc.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def write_traffic(fetch_csv: str, write_csv: str, copy_fetch_csv: str, copy_write_csv: str, key: str) -> dict:
    """profiles/traffic.json from two rocprofv3 --pmc runs of the benchmark
    (FETCH_SIZE, WRITE_SIZE, KiB per dispatch) and the same two counters for
    a plain 1 GiB device copy (bench.py --calibrate-copy), which scale the
    counters to bytes; stamped with the kernel sources they were measured
    at."""
    gib = float(1 << 30)

    def copy_factor(path, counter):
        vals = [v for k, v in counter_means(path, counter).items() if "copyBuffer" in k]
        if not vals:
            raise SystemExit(f"no copy kernel in {path}")
        return gib / (max(vals) * 1024.0)

    fr, fw = copy_factor(copy_fetch_csv, "FETCH_SIZE"), copy_factor(copy_write_csv, "WRITE_SIZE")
    fetch, write = counter_means(fetch_csv, "FETCH_SIZE"), counter_means(write_csv, "WRITE_SIZE")
    out = {}
    for name, prefix in TRAFFIC_KERNELS.items():
        kf = [k for k in fetch if k.startswith(prefix)]
        kw = [k for k in write if k.startswith(prefix)]
        if kf and kw:
            out[f"{key}:{name}"] = int(round(fetch[kf[0]] * 1024.0 * fr + write[kw[0]] * 1024.0 * fw))
    data = {"kernel_sources_sha256": kernel_source_digest(), "read_scale": fr, "write_scale": fw,
            "bytes_per_launch": out}
    with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as f:
        json.dump(data, f, indent=1)
        f.write("\n")
    return data


def cpu_baseline(spec, budget_s: float, seed: int, ens_workers: int, sim_steps: int):
    """CPU figures beside the GPU ones, all from this project's C restatements
    of the reference's CPU solvers:
    * value: the NSM port on the benchmark model, one core, bounded by wall
      time; equivalent site-steps/s = N x (simulated time / tau) / wall;
    * all_cores_ensemble: the ensemble workload's replicates run by the NSM
      port one replicate per host core (the reference's ReplicateRunner runs
      one replicate per thread), as simulations per hour;
    * serial_mpd_oracle: one step of the serial MPD restatement of the
      benchmark model, for context."""
    import threading

    import oracle
    from lattice_microbes_amd import models

    sim_t, events, wall, setup_s = nsm_run(spec, seed, 0, budget_s)
    eq_steps = sim_t / spec.tau
    out = {
        "value": spec.N * eq_steps / wall,
        "unit": "site-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"next-subvolume port of NextSubvolumeSolver on the same {spec.name} model "
                   f"({spec.N} subvolumes): {events} events in {wall:.1f} s wall = "
                   f"{sim_t * 1e6:.3f} us simulated = {eq_steps:.3f} tau-steps equivalent "
                   f"(setup {setup_s:.1f} s untimed)"),
        "events_per_s": events / wall,
        "host": host_cpu(),
    }
    # all cores: replicate r of the ensemble mix on thread r
    mix = models.ensemble_mix(max(ens_workers, 1), seed)
    res = [None] * len(mix)

    def work(k):
        spec_k, r, sd = mix[k]
        res[k] = nsm_run(spec_k, sd, r, budget_s)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(mix))]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    span = time.perf_counter() - t0
    if any(r is None for r in res):
        raise RuntimeError("an NSM ensemble worker failed")
    tau = mix[0][0].tau
    sims_per_s = sum(r[0] / (sim_steps * tau) / r[2] for r in res)
    out["all_cores_ensemble"] = {
        "workload": f"bimolecular replicates of the ensemble mix ({', '.join(str(m[0].shape[0]) for m in mix)} cubed)",
        "threads": len(mix), "sim_steps": sim_steps,
        "sims_per_hr": sims_per_s * 3600.0,
        "events_per_s": sum(r[1] / r[2] for r in res),
        "wall_s": span,
    }
    # the serial MPD restatement (the oracle), one step of the benchmark model
    orc = oracle.Oracle(spec, seed_top=seed)
    t0 = time.perf_counter()
    if orc.run(1) != 0:
        raise RuntimeError("oracle step failed")
    step_s = time.perf_counter() - t0
    out["serial_mpd_oracle"] = {"s_per_step": step_s, "site_steps_per_s": spec.N / step_s, "cores": 1}
    return out


def ensemble_bench(nrep: int, steps: int, seed: int, sync_steps: int):
    """Batched ensemble: nrep replicates of the bimolecular model on uneven
    lattices (32^3, 48^3, 64^3, 96^3, 128^3 drawn uniformly), each with its own
    seed, stepped by one Ensemble, against plain MPD one simulation at a time:
    each replicate runs all its steps (launches queued back to back) before
    the next one starts; and, as the reference steps (a host synchronisation
    after every step, MpdRdmeSolver.cu:1091), sync_steps steps of every
    replicate.  Every replicate's final lattice is compared between the
    batched and the one-at-a-time runs."""
    import hashlib

    import lattice_microbes_amd as lmx
    from lattice_microbes_amd import models

    mix = models.ensemble_mix(nrep, seed)
    sites = sum(spec.N for spec, _, _ in mix)

    def make(spec, r, sd):
        s = lmx.MpdRdmeSolver(0)
        spec.load_into(s, replicate=r, seed=sd)
        return s

    def digest(s):
        return hashlib.sha256(s.getLatticeNative().tobytes()).hexdigest()

    # batched: one warm step (graph capture), then the timed steps
    batch = [make(*m) for m in mix]
    ens = lmx.Ensemble(0)
    for s in batch:
        ens.add(s)
    ens.run(1)
    tuned = ens.autotune(steps=8)  # launch configuration from the batch autotuner (state restored)
    ms = ens.time_steps(steps)
    batched = [digest(s) for s in batch]
    ens.close()
    for s in batch:
        s.close()
    del batch
    # plain, one simulation at a time: 1 + steps steps each, the timed steps
    # measured with HIP events on the solver's stream
    plain_ms = 0.0
    same = 0
    for (spec, r, sd), d in zip(mix, batched):
        s = make(spec, r, sd)
        s.runTimesteps(1)
        t, _ = s.timeSteps(steps)
        plain_ms += t
        same += int(digest(s) == d)
        s.close()
    # the reference's cadence: the host waits for every step
    sync_wall = 0.0
    for spec, r, sd in mix:
        s = make(spec, r, sd)
        s.runTimesteps(1)
        t0 = time.perf_counter()
        for _ in range(sync_steps):
            s.runTimesteps(1)
        sync_wall += time.perf_counter() - t0
        s.close()
    # one persistent fused diffuse+react launch per step against the sweeps,
    # on a single replicate of each lattice size: a batch of fused steps does
    # the same per-site work, so it can only win where the fused kernel does
    paths = {}
    for L in models.ENSEMBLE_SIZES:
        spec = next(m[0] for m in mix if m[0].shape[0] == L)
        row = {}
        for name, path in (("sweeps", lmx._lib.PATH_SWEEPS), ("fused", lmx._lib.PATH_FUSED)):
            s = make(spec, 0, 7)
            s.setPath(path)
            s.runTimesteps(2)
            t, _ = s.timeSteps(200)
            row[name + "_us_per_step"] = t / 200 * 1e3
            s.close()
        paths[f"{L}^3"] = row
    b_rate = sites * steps / (ms / 1e3)
    p_rate = sites * steps / (plain_ms / 1e3)
    sync_rate = sites * sync_steps / sync_wall
    sim_steps = 20000
    return {
        "workload": f"{nrep} bimolecular replicates, lattices drawn from {list(models.ENSEMBLE_SIZES)} cubed, "
                    "own seeds",
        "replicates": nrep, "sites": sites, "steps_timed": steps,
        "batched_site_steps_per_s": b_rate,
        "plain_one_at_a_time_site_steps_per_s": p_rate,
        "plain_stepsync_site_steps_per_s": sync_rate, "stepsync_steps_timed": sync_steps,
        "speedup_vs_plain": b_rate / p_rate, "speedup_vs_plain_stepsync": b_rate / sync_rate,
        "sims_per_hr_batched": nrep / (ms / 1e3 / steps * sim_steps) * 3600.0,
        "sims_per_hr_plain": nrep / (plain_ms / 1e3 / steps * sim_steps) * 3600.0,
        "sims_per_hr_plain_stepsync": nrep / (sync_wall / sync_steps * sim_steps) * 3600.0,
        "sim_steps": sim_steps,
        "bit_exact_vs_plain": {"checked": nrep, "identical": same},
        "fused_vs_sweeps_single_replicate": paths,
        "autotune": tuned,
    }


def uint32_bench(spec, steps: int, seed: int, copy_gbps: float):
    """The same model on the uint32 lattice (IntMpdRdmeSolver) with every
    species id raised by 300 past a byte: per-step time and HBM rate of its
    one-site-per-thread kernels (2*4P+1 B/site per sweep, 4P+1 B/site for the
    reaction phase)."""
    import lattice_microbes_amd as lmx
    from lattice_microbes_amd.models import ModelSpec

    off = 300
    nS = spec.nS + off
    S = np.zeros((nS, spec.R), np.int32)
    S[off:] = spec.S
    D = np.zeros((nS, spec.R), np.uint32)
    D[off:] = spec.D
    DF = np.zeros((spec.nST, spec.nST, nS))
    DF[..., off:] = spec.DF
    lat = spec.lattice.astype(np.uint32)
    lat[lat > 0] += off
    wspec = ModelSpec(name=spec.name, shape=spec.shape, spacing=spec.spacing, tau=spec.tau, P=spec.P, nS=nS,
                      nST=spec.nST, types=spec.types, K=spec.K, S=S, D=D, DF=DF, RL=spec.RL, lattice=lat,
                      sites=spec.sites, boundary=spec.boundary)
    del lat
    w = lmx.IntMpdRdmeSolver(0)
    try:
        wspec.load_into(w, replicate=0, seed=seed)
        w.runTimesteps(3)
        total_ms, _ = w.timeSteps(steps)                  # events around the loop only
        _, per = w.timeSteps(steps, per_kernel=True)      # per-launch events for the kernel averages
    finally:
        w.close()
    N, P = spec.N, spec.P
    sweep_b, react_b = (8 * P + 1) * N, (4 * P + 1) * N
    step_b = 3 * sweep_b + (react_b if spec.R else 0)
    ms = total_ms / steps
    gbps = step_b / (ms / 1e3) / 1e9
    return {"workload": f"{spec.name} {spec.shape[0]}x{spec.shape[1]}x{spec.shape[2]}, species ids +{off} "
                        f"({nS} species), 4 bytes per particle",
            "steps": steps, "ms_per_step": ms, "site_steps_per_s": N * steps / (total_ms / 1e3),
            "kernels_us": {k: v[0] / max(v[1], 1) * 1e3 for k, v in per.items()},
            "alg_bytes_per_step": step_b, "achieved_GBps": gbps, "frac": gbps / HBM_PEAK_GBS,
            "frac_of_copy_peak": gbps / copy_gbps if copy_gbps else None}


def serving_bench(njobs: int, seed: int):
    """Continuous batching of a Poisson request stream (20 req/s) mixing the
    bimolecular 64^3, capsule 128x128x256 and whole-cell 256x256x128 (30
    species) models, 1e3-1e4 steps each, admitted smallest work first (a job
    waiting over 5 s goes first) while the batch holds at most 32M sites,
    against the same requests served one at a time in arrival order; every
    job's final lattice is compared between the two.  The reference's cadence
    (a host synchronisation after every step) is timed on every fourth
    request."""
    from lattice_microbes_amd import models, serving

    templates = models.serving_templates(seed)
    mk = lambda: serving.poisson_jobs(templates, njobs, 20.0, (1000, 10000), seed=seed + 10)
    jobs = mk()
    served = serving.Server(0, hbm_budget=200 << 30, max_sites=32 << 20, policy="sjf").run(jobs)
    digest = {j.id: j.digest for j in jobs}
    jobs2 = mk()
    plain = serving.serve_plain(jobs2)
    same = sum(int(digest[j.id] == j.digest) for j in jobs2)
    sample = mk()[::4]
    for j in sample:
        j.arrival = 0.0
    sync = serving.serve_plain(sample, step_sync=True, keep_digest=False)
    return {
        "workload": f"{njobs} requests, Poisson 20/s, bimolecular 64^3 / capsule 128x128x256 / "
                    "whole-cell 256x256x128 (30 species), 1e3-1e4 steps",
        "admission": "smallest work first (sites x steps), a request waiting over 5 s first; <= 32M sites per step",
        "served": served, "plain": plain,
        "throughput_vs_plain": served["site_steps_per_s"] / plain["site_steps_per_s"],
        "plain_stepsync_sample": {"jobs": sync["jobs"], "site_steps_per_s": sync["site_steps_per_s"]},
        "throughput_vs_plain_stepsync": served["site_steps_per_s"] / sync["site_steps_per_s"],
        "p50_vs_plain": served["p50_latency_s"] / plain["p50_latency_s"],
        "p99_vs_plain": served["p99_latency_s"] / plain["p99_latency_s"],
        "bit_exact_vs_plain": {"jobs": njobs, "identical": same},
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
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-uint32", action="store_true", help="skip the uint32-lattice measurement")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default kernels instead of timing every y/z kernel and window shape first")
    ap.add_argument("--ensemble-replicates", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=2000)
    ap.add_argument("--ensemble-sync-steps", type=int, default=100,
                    help="steps per replicate of the per-step-synchronised plain run")
    ap.add_argument("--serving-jobs", type=int, default=100)
    ap.add_argument("--write-traffic", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="write profiles/traffic.json from rocprofv3 FETCH_SIZE/WRITE_SIZE runs of the benchmark "
                         "and of --calibrate-copy, then exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()

    if args.write_traffic:
        spec_key = {"whole_cell": "whole_cell_256x256x512"}.get(args.workload, args.workload)
        print(json.dumps(write_traffic(*args.write_traffic, spec_key)), flush=True)
        return 0

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
        print("bench.py runs the MPD-RDME step on one GPU only: sharding a lattice into z slabs over "
              "several GPUs (one exchange of boundary planes per step) is designed but not built.",
              file=sys.stderr)
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
    tuned = None
    if args.path == "sweeps" and not args.no_autotune:
        tuned = solver.autotune(steps=8)
        log(f"autotune: {tuned}")
    if args.warmup:
        solver.runTimesteps(args.warmup)
    log(f"warmup {args.warmup} steps done")
    # The timed region: K steps back to back, HIP events around the loop only.
    total_ms, _ = solver.timeSteps(args.steps)
    log(f"timed {args.steps} steps: {total_ms:.2f} ms")
    # Per-kernel durations from a second run of K steps with HIP events around
    # every launch (the event records add gaps between kernels, so this run's
    # total is not the headline).
    evented_ms, per = solver.timeSteps(args.steps, per_kernel=True)
    log(f"per-kernel run of {args.steps} steps: {evented_ms:.2f} ms")
    counts = solver.getSpeciesCounts()
    ovf_entries, ovf_steps = solver.getOverflowStats()
    native = solver.getLatticeNative()
    occupied = int(np.count_nonzero(np.bitwise_or.reduce(native, axis=0)))  # sites holding particles

    N, P, K = spec.N, spec.P, args.steps
    value = N * K / (total_ms / 1e3)
    tile_list = tuned is None or (tuned["path"] == 0 and tuned["yz_kernel"] == 0 and tuned["fuse_reaction"])
    listed = args.path == "sweeps" and uses_reaction_list(spec) and tile_list
    path = "fused" if args.path == "fused" or (tuned and tuned["path"] == 1) else "sweeps"
    list_bytes = occupied * list_entry_bytes(P) if listed else 0
    # per-kernel averages and algorithmic bytes per launch (with the reaction
    # list, the z sweep also writes one entry per occupied site and the
    # reaction kernel reads them)
    kernels = {}
    for name, (ms, launches) in per.items():
        if name.startswith("sweep") or name == "fused":
            b = sweep_bytes_per_site(P) * N + (list_bytes if name == "sweep_z" else 0)
        elif name == "reaction":
            b = list_bytes if listed else reaction_bytes_per_site(P) * N
        else:
            b = 0
        avg_ms = ms / max(launches, 1)
        kernels[name] = {"avg_us": avg_ms * 1e3, "launches": launches,
                         "alg_bytes": b, "GBps": (b / (avg_ms / 1e3) / 1e9) if b and avg_ms > 0 else None}
    dom = max((k for k in kernels if kernels[k]["alg_bytes"]), key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    dk = kernels[dom]
    traffic, traffic_note = measured_traffic(f"{spec.name}_{spec.shape[0]}x{spec.shape[1]}x{spec.shape[2]}:{dom}")
    # Measured copy peak (1 GiB copy, one 16-byte load and store per thread;
    # read + write bytes).
    copy_gbps = ctypes.c_double(0.0)
    lmx._lib.check(lmx.load_library().lmx_copy_bandwidth(0, 1 << 30, 20, ctypes.byref(copy_gbps)))
    roofline = {"bound": "hbm", "kernel": dom, "achieved": dk["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": dk["GBps"] / HBM_PEAK_GBS, "traffic": traffic,
                "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_us": dk["avg_us"],
                "copy_peak_measured": copy_gbps.value, "frac_of_copy_peak": dk["GBps"] / copy_gbps.value,
                # HBM bytes actually moved (rocprofv3 FETCH/WRITE) per launch time: below the
                # algorithmic rate where the kernels skip the zero upper planes of packed sites
                "traffic_GBps": (traffic / (dk["avg_us"] / 1e6) / 1e9) if traffic else None,
                "traffic_source": traffic_note}
    # Bytes of the step as run: three sweeps plus the reaction list (written
    # and read once), or plus the full reaction read.
    unfused_alg = (3 * sweep_bytes_per_site(P) + (reaction_bytes_per_site(P) if spec.R else 0)) * N
    step_alg = 3 * sweep_bytes_per_site(P) * N + 2 * list_bytes if listed else unfused_alg
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
                   "lambda_m": spec.spacing, "tau_s": spec.tau, "path": path,
                   "reaction_phase": "list" if listed else "full pass", "occupied_sites_end": occupied},
        "roofline": roofline,
        "step_roofline": {"alg_bytes_per_step": step_alg, "achieved_GBps": step_alg / (step_ms / 1e3) / 1e9,
                          "frac": step_alg / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                          "unfused_equivalent_bytes_per_step": unfused_alg,
                          "unfused_equivalent_frac": unfused_alg / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                          # a step that read and wrote the lattice once (2P+1 B/site)
                          "fused_floor_bytes_per_step": sweep_bytes_per_site(P) * N,
                          "fused_floor_frac": sweep_bytes_per_site(P) * N / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS},
        "kernels": kernels,
        "overflow": {"entries": ovf_entries, "steps_with_overflow": ovf_steps},
        "final_particles": int(counts.sum()),
        "ms_per_step_with_kernel_events": evented_ms / K,
        # launch shape chosen by lmx_autotune (8 timed steps per candidate) before the warmup
        "autotune": tuned,
    }
    solver.close()
    if not args.no_uint32 and args.path == "sweeps":
        log("uint32 lattice")
        out["uint32_lattice"] = uint32_bench(spec, 30, args.seed, copy_gbps.value)
    if not args.no_ensemble:
        log(f"ensemble: {args.ensemble_replicates} replicates")
        out["ensemble"] = ensemble_bench(args.ensemble_replicates, args.ensemble_steps, args.seed,
                                         args.ensemble_sync_steps)
    if not args.no_serving:
        log(f"serving: {args.serving_jobs} requests")
        out["serving"] = serving_bench(args.serving_jobs, args.seed)
    if not args.no_cpu:
        log(f"cpu baseline ({args.cpu_budget:.0f}s budget)")
        out["cpu_baseline"] = cpu_baseline(spec, args.cpu_budget, args.seed, min(16, host_cpu()["affinity"] or 1),
                                           20000)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
