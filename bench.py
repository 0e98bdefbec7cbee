"""Benchmark: particle-steps/s of the DSMC step on one MI355X.

Workload (default, c3): the large single-case configuration - 2000 x 2000 grid on
[-20,20]^2 m, 4x4 m box, NPCFree 20 (~8e7 particles), Ma 4, n 1e17 m^-3, T 200 K, AoA -5 deg,
CLL walls (acc_n = acc_t = 0.8) at 300 K, Dt 1e-5 s, 5000 steps sampling after step 2000.
Synthetic data: the freestream initial fill, developed by this path's own steps up to the timed
window. The window sits where the run's own share of sampling steps holds (3000 of 5000 steps
sample, so 12 of 20 timed steps do); W warmup steps precede it. value = sum over the K timed
steps of the particles indexed that step / wall-clock of the K fused, graph-captured steps.

Before the development the autotuner picks the launch schedule and gather chunk (speed only;
its steps count as development). After the timed region, eager steps whose kernels carry
start/stop events from their own dispatch give each kernel's mean duration, on non-sampling and
sampling steps weighted by the timed region's mix of the two; the dominant kernel's algorithmic
bytes over that duration is the roofline figure, and its HBM traffic comes from the committed
rocprofv3 PMC summary. The CPU baseline runs the reference's own program under mpiexec on every
host core this process may use (oracle/_ref, compiled from its sources) at the 100 x 100 case,
the largest its fixed per-rank arrays hold, beside the oracle's restatement of the reference's
step functions on one core at the workload itself.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (nx, ny, npc, wall, aoa, first_sampling, total_steps, extra)
    "c2": dict(nx=100, ny=100, npc_free=10.0, wall_model=0, attack_angle_deg=15.0,
               first_sampling_step=10000, n_step=20000,
               label="100x100 grid, NPCFree 10, Ma 4, AoA 15, diffuse Tw 300 K, 20000 steps"),
    "c3": dict(nx=2000, ny=2000, npc_free=20.0, wall_model=2, acc_n=0.8, acc_t=0.8,
               attack_angle_deg=-5.0, first_sampling_step=2000, n_step=5000,
               label="2000x2000 grid, NPCFree 20, Ma 4, AoA -5, CLL 0.8/0.8 Tw 300 K, 5000 steps"),
    "c1": dict(nx=50, ny=50, npc_free=10.0, wall_model=1, attack_angle_deg=0.0,
               first_sampling_step=1000, n_step=3000,
               label="50x50 grid, NPCFree 10, Ma 4, AoA 0, specular, 3000 steps"),
    # the batched sweep: 64 cases drawn from seed 0 (dsmc_erlend_amd.sweep.draw_cases) in one
    # particle pool
    "c4": dict(batch=64, seed=0, n_step=5000, first_sampling_step=2000,
               label="64-case batched sweep from seed 0 (AoA [-30,30], walls cycled, grid side "
                     "[200,800], NPCFree [5,40], Ma [2,8], n [1e16,1e18]), 5000 steps, sampling "
                     "after step 2000"),
}


def sampling_share(wl):
    """Share of a workload's steps that sample: Step > FirstSamplingStep of NStep."""
    return max(0, wl["n_step"] - wl["first_sampling_step"] - 1) / wl["n_step"]


def host_cores():
    """Cores this process may run on."""
    return len(os.sched_getaffinity(0))


def stage_bytes(stage, n, m, sampling):
    """Algorithmic HBM bytes of one launch of a stage's kernel: the per-unit bytes of the
    reference step functions the kernel carries out, every array element read or written once
    (FP64 x,y and FP32 vx,vy,vz; u32 keys, offsets, positions). The gather is Indexing's stable
    scatter (R pool position 4 + R particle 28 + W particle 28 per particle), CollideParticles'
    per-cell reads (offsets 8, (sigma*g)max R/W 8) and, on sampling steps, Sampling (R velocity
    12 per particle; the 17-double CELL read and written, 272, and the cell's count 8)."""
    if stage == "move":      # R x,y 16 + vx,vy 8; W key 4; counts R/W 8 per cell (the fused
        # step's gather writes the moved positions as it copies them)
        return 28 * n + 8 * m
    if stage == "scatter":   # R key 4, W perm 4 per particle; cursors R/W 8 per cell
        return 8 * n + 8 * m
    if stage == "gather":
        return 60 * n + 16 * m + ((12 * n + 280 * m) if sampling else 0)
    if stage == "scan":      # counts R, offsets and cursors W, counts cleared
        return 16 * m
    return 0


def step_bytes(n, m, n_in, ncoll, sampling):
    """Algorithmic bytes of a whole step of the reference's functions, each array element read
    or written once: MoveParticles R/W 56, Indexing's key and histogram 20 and stable scatter 60
    per particle, CollideParticles 16 per cell and 48 per collision, the scan 8 per cell, the
    new particles 28 each; Sampling adds 12 per particle and 280 per cell on sampling steps."""
    b = 136 * n + 28 * m + 28 * n_in + 48 * ncoll
    return b + (12 * n + 280 * m if sampling else 0)


def kernel_bytes(stage, n, m, sampling):
    """The bytes the kernel itself needs where that is less than stage_bytes: the fused gather
    samples from its LDS copy of the velocities (no second read) and keeps 13 FP64 sums per
    cell (the reference's symmetric CountVV entries once), 208 B/cell read and written."""
    if stage == "gather":
        return 60 * n + 16 * m + (208 * m if sampling else 0)
    return stage_bytes(stage, n, m, sampling)


def profiled_traffic(workload, stage, frac_samp):
    """HBM bytes per launch of a stage's kernel from the committed rocprofv3 PMC summaries
    (profiles/*.md lines `traffic_bytes_per_launch <workload> <stage> <bytes> [plain|sampling]`:
    FETCH_SIZE and WRITE_SIZE runs scaled by a copy of known size, on steps without or with
    moment sampling; a line without a kind covers both), weighted by the timed region's share
    `frac_samp` of sampling steps; (None, None) when a needed kind was not recorded."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None, None
    found = {}
    for name in sorted(os.listdir(d)):
        if not name.endswith(".md"):
            continue
        for line in open(os.path.join(d, name)):
            f = line.split()
            if len(f) in (4, 5) and f[0] == "traffic_bytes_per_launch" and f[1] == workload \
                    and f[2] == stage:
                found.setdefault(f[4] if len(f) == 5 else "both", (float(f[3]), f"profiles/{name}"))
    if "both" in found:
        return found["both"]
    weights = [(k, w) for k, w in (("plain", 1.0 - frac_samp), ("sampling", frac_samp)) if w > 0]
    if not weights or any(k not in found for k, _ in weights):
        return None, None
    value = sum(found[k][0] * w for k, w in weights)
    return value, ", ".join(sorted({found[k][1] for k, _ in weights}))


def calibrate():
    """A plain elementwise copy of known size (1 GiB read + 1 GiB written per launch) for
    scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "float32 copy", "bytes_read": a.numel() * 4,
                      "bytes_written": b.numel() * 4, "launches": 5}))


def fill_particles(wl):
    """Particles of the workload's freestream fill (NX * NY * NPCFree, box area ignored)."""
    return wl["nx"] * wl["ny"] * wl["npc_free"]


def cpu_sample_steps(wl, budget, cap):
    """Steps of a CPU sample of about `budget` particle-steps: `cap` for the 100 x 100 case,
    a few for a 2000 x 2000 grid."""
    return max(2, min(cap, int(budget / fill_particles(wl))))


# The reference's fixed per-rank arrays (include/Constants.h:4-7): CELL C[200][200], PCL P[1e6].
REF_MAX_CELLS_PER_AXIS = 200
REF_MAX_PARTICLES_PER_RANK = 1000000


def mpi_process_grid(ranks):
    """(ProcX, ProcY) of the reference's SetupMPIDomain (src/MPIManager.cpp:85-108):
    MPI_Dims_create's most balanced factorisation, the larger factor along x."""
    py = max(a for a in range(1, int(ranks ** 0.5) + 1) if ranks % a == 0)
    return ranks // py, py


def mpi_fits(wl, ranks):
    """Whether each rank's subdomain fits the reference's fixed arrays (with a 1.5x margin on
    the freestream particle count for the denser gas ahead of the box)."""
    px, py = mpi_process_grid(ranks)
    return (-(-wl["nx"] // px) <= REF_MAX_CELLS_PER_AXIS
            and -(-wl["ny"] // py) <= REF_MAX_CELLS_PER_AXIS
            and 1.5 * fill_particles(wl) / ranks <= REF_MAX_PARTICLES_PER_RANK)


def mpi_baseline(wl, ranks, share, seconds=10.0):
    """The reference's own program (RunSimulation with its MPI domain decomposition,
    oracle/_ref/ref_mpi_baseline) on `ranks` host cores: two runs of s1 and s2 steps differenced,
    so setup and output drop out; the particles of steps s1..s2 come from the reference's own
    progress lines. The span s2 - s1 is sized to about `seconds` of the longer run at an assumed
    3e7 particle-steps/s per rank, and its last `share` samples (FirstSamplingStep placed so), as
    the GPU's timed window does."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_mpi_baseline")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not os.path.exists(exe) or not os.path.exists(mpiexec):
        return None
    if not mpi_fits(wl, ranks):
        print(f"[bench] cpu baseline: a {wl['nx']} x {wl['ny']} grid does not fit the "
              f"reference's per-rank arrays on {ranks} ranks", file=sys.stderr, flush=True)
        return None
    span = max(200, int(seconds * 3e7 * ranks / fill_particles(wl)) // 100 * 100)
    s1 = 1000
    s2 = s1 + span
    first = s2 - int(round(share * span))
    args = [str(wl["nx"]), str(wl["ny"]), str(wl["npc_free"]), str(wl["wall_model"]),
            str(wl["attack_angle_deg"])]
    runs = {}
    try:
        for steps in (s1, s2):
            print(f"[bench] cpu baseline: the reference under mpiexec -n {ranks}, {steps} steps",
                  file=sys.stderr, flush=True)
            r = subprocess.run([mpiexec, "-n", str(ranks), exe] + args + [str(steps), str(first)],
                               capture_output=True, text=True, timeout=600,
                               cwd=os.path.join(ROOT, "oracle", "_ref"))
            if r.returncode != 0:
                print(f"mpi baseline failed: {r.stderr[-400:]}", file=sys.stderr)
                return None
            d = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
            nps = {}
            for line in r.stdout.splitlines():
                if line.startswith("Time step"):
                    f = line.replace(",", "").split()
                    nps[int(f[2])] = int(f[6])
            runs[steps] = (d["seconds"], nps)
    except Exception as e:  # noqa: BLE001
        print(f"mpi baseline failed: {e}", file=sys.stderr)
        return None
    (ta, _), (tb, nps) = runs[s1], runs[s2]
    counted = [n for st, n in nps.items() if s1 < st <= s2]
    if tb <= ta or not counted:
        return None
    value = sum(counted) / len(counted) * (s2 - s1) / (tb - ta)
    return {"value": value, "unit": "particle-steps/s", "cores": ranks, "kind": "reference",
            "sample": f"the reference's RunSimulation under mpiexec -n {ranks} (every core this "
                      f"process may use) on the {wl['nx']} x {wl['ny']} case (NPCFree "
                      f"{wl['npc_free']:g}, the largest its fixed arrays hold: the workload's grid "
                      f"does not fit them); runs of {s1} and {s2} steps from the freestream fill "
                      f"(sampling after step {first}) differenced: {tb - ta:.2f} s for steps "
                      f"{s1}..{s2}"}


def cpu_baseline(wl, steps, first_sampling):
    """Times the reference's own functions (kind 'reference') or, failing that, the oracle's
    restatement (kind 'port') on one core, on a bounded sample of the workload."""
    args = [str(wl["nx"]), str(wl["ny"]), str(wl["npc_free"]), str(wl["wall_model"]),
            str(wl["attack_angle_deg"]), str(steps)]
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_baseline")
    port = os.path.join(ROOT, "oracle", "cpu_baseline")
    tries = []
    if os.path.exists(ref):
        tries.append(("reference", [ref] + args + [str(first_sampling)]))
    if os.path.exists(port):
        tries.append(("port", [port] + args + ["0", str(first_sampling)]))
    for kind, cmd in tries:
        try:
            print(f"[bench] cpu baseline: one core ({kind}), {steps} steps", file=sys.stderr,
                  flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
            if r.returncode == 0:
                line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
                d = json.loads(line)
                return {"value": d["particle_steps_per_s"], "unit": "particle-steps/s",
                        "cores": d["cores"], "kind": kind,
                        "sample": f"steps 0..{steps - 1} from the freestream fill, sampling after "
                                  f"step {first_sampling}, {d['seconds']:.1f} s on one core"}
        except Exception as e:  # noqa: BLE001
            print(f"cpu baseline {kind} failed: {e}", file=sys.stderr)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="steps of the one-core sample (default: about 2.4e8 particle-steps, "
                         "at most 2000)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tuning", default=None,
                    help="schedule,chunk[,particles_per_thread] to set instead of autotuning "
                         "(schedule small|large)")
    ap.add_argument("--autotune-steps", type=int, default=20,
                    help="fused steps per candidate of dsmc_autotune before the development (0: "
                         "the default tuning)")
    ap.add_argument("--eager-develop", action="store_true",
                    help="develop the fill with eager steps (kernels launched one by one, as a "
                         "profiler traces them) instead of graph replays")
    ap.add_argument("--fresh", action="store_true",
                    help="time the freestream fill itself: no development before the warmup (the "
                         "step counter is still set so the window keeps the sampling share)")
    ap.add_argument("--save-state", default=None,
                    help="after the timed region, save the pool and step counter (one case) to "
                         "this .npz, so a profiler can time short runs of the developed flow")
    ap.add_argument("--load-state", default=None,
                    help="replace the pool by one saved with --save-state before the warmup (one "
                         "case; no development)")
    ap.add_argument("--start-step", type=int, default=None,
                    help="set the step counter before the warmup instead of placing the window "
                         "by the sampling share (for per-launch counter runs)")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only a copy kernel of known size (for PMC byte counters)")
    a = ap.parse_args()
    if a.calibrate:
        calibrate()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the path would "
              "shard by whole cases, one particle pool per GPU)", file=sys.stderr)
        sys.exit(2)

    from dsmc_erlend_amd import Simulation, default_params
    wl = WORKLOADS[a.workload]
    spec = dict(wl)
    label = spec.pop("label")
    total = spec["n_step"]
    first = spec["first_sampling_step"]
    steps = a.steps if a.steps is not None else 200
    warmup = a.warmup if a.warmup is not None else 10
    if "batch" in spec:
        from dsmc_erlend_amd.sweep import draw_cases
        cases = [q.replace(first_sampling_step=first, n_step=total)
                 for q in draw_cases(spec["batch"], spec["seed"])]
    else:
        cases = [default_params(**spec)]
    p = cases[0]
    sim = Simulation(cases)
    nc = len(cases)
    t_init = time.perf_counter()
    sim.initial_conditions()
    t_init = time.perf_counter() - t_init

    def pool():
        return sum(sim.num_particles(i) for i in range(nc))

    def psteps_now():
        return sum(sim.state(i).particle_steps for i in range(nc))

    print(f"[bench] {a.workload}: {nc} case(s), {pool()} particles after init "
          f"({t_init:.2f} s)", file=sys.stderr, flush=True)
    tuning = sim.tuning()
    if a.tuning:
        sched, chunk, *ppt = a.tuning.split(",")
        sim.set_tuning(schedule=1 if sched == "small" else 0, chunk=int(chunk),
                       particles_per_thread=int(ppt[0]) if ppt else 1)
        tuning = sim.tuning()
    elif a.autotune_steps > 0:
        tuning = sim.autotune(a.autotune_steps)   # speed only: results do not depend on it
        print(f"[bench] autotuned (schedule, chunk, particles per thread) = {tuning}",
              file=sys.stderr, flush=True)
    # the timed window [t_start, t_start + steps) holds the run's own share of sampling steps
    share = sampling_share(wl)
    n_samp_target = int(round(share * steps))
    t_start = max(0, first + 1 + n_samp_target - steps)
    developed = 0
    if a.load_state:
        if nc != 1:
            sys.exit("--load-state takes a single-case workload")
        saved = np.load(a.load_state)
        sim.set_particles(saved["x"], saved["y"], saved["vx"], saved["vy"], saved["vz"])
        sim.set_step(int(saved["step"]))
        print(f"[bench] loaded {saved['x'].size} particles at step {int(saved['step'])} from "
              f"{a.load_state}", file=sys.stderr, flush=True)
    if a.start_step is not None:
        sim.set_step(a.start_step)
    elif a.fresh or a.load_state:
        sim.set_step(max(0, t_start - warmup))
    else:
        # develop the fill with this path's own steps up to the warmup
        todo = max(0, t_start - warmup - sim.state(0).step)
        t_dev = time.perf_counter()
        if a.eager_develop:
            for done in range(0, todo, 250):
                sim.profile(min(250, todo - done))
        else:
            sim.run(todo)
        developed = sim.state(0).step
        print(f"[bench] developed the fill to step {developed} ({todo} steps, "
              f"{time.perf_counter() - t_dev:.1f} s), {pool()} particles", file=sys.stderr,
              flush=True)
    if warmup:
        sim.run(warmup)
    ps0 = psteps_now()
    step0 = sim.state(0).step
    coll0 = sum(sim.state(i).total_collisions for i in range(nc))
    t0 = time.perf_counter()
    sim.run(steps)              # synchronises at the end
    t1 = time.perf_counter()
    psteps = psteps_now() - ps0
    coll_per_step = (sum(sim.state(i).total_collisions for i in range(nc)) - coll0) / steps
    inj_per_step = sum(sim.state(i).last_injected for i in range(nc))
    np_end = pool()
    wall = t1 - t0
    value = psteps / wall
    print(f"[bench] timed {steps} steps in {wall:.3f} s: {value:.4e} particle-steps/s, "
          f"{np_end} particles", file=sys.stderr, flush=True)
    if a.save_state:
        if nc != 1:
            sys.exit("--save-state takes a single-case workload")
        x, y, vx, vy, vz = sim.particles(0)
        try:
            np.savez(a.save_state, x=x, y=y, vx=vx, vy=vy, vz=vz, step=sim.state(0).step)
        except OSError as e:
            print(f"[bench] could not save the state: {e}", file=sys.stderr, flush=True)
        del x, y, vx, vy, vz

    # Per-kernel times: eager steps whose kernels carry start/stop events from their own
    # dispatch, taken on the developed pool on steps of both kinds the timed region held
    # (sampling steps add the moment sums to the gather), weighted by that region's mix.
    n_samp = max(0, (step0 + steps - 1) - max(step0, first + 1) + 1)
    frac_samp = n_samp / steps
    n = pool()
    m = sum(q.nx * q.ny for q in cases)
    profs = []
    if frac_samp < 1.0:
        sim.set_step(max(0, min(step0, first - a.profile_steps)))
        profs.append((1.0 - frac_samp, sim.profile(a.profile_steps), False))
    if frac_samp > 0.0:
        sim.set_step(max(first + 1, step0))
        profs.append((frac_samp, sim.profile(a.profile_steps), True))
    prof = {k: sum(w * pr[k] for w, pr, _ in profs) for k in profs[0][1]}
    # single-kernel stages (the scan is three launches on large grids)
    kernels = {k: v for k, v in prof.items() if k in ("move", "gather", "scatter")}
    dom = max(kernels, key=kernels.get)
    ms = prof[dom]
    nbytes = sum(w * stage_bytes(dom, n, m, smp) for w, _, smp in profs)
    achieved = nbytes / (ms * 1e-3) / 1e9
    kbytes = sum(w * kernel_bytes(dom, n, m, smp) for w, _, smp in profs)
    # the whole step's algorithmic bytes over the timed region's mean step (its own mix of
    # sampling and plain steps)
    sbytes = ((1.0 - frac_samp) * step_bytes(psteps / steps, m, inj_per_step, coll_per_step, False)
              + frac_samp * step_bytes(psteps / steps, m, inj_per_step, coll_per_step, True))
    traffic, traffic_src = profiled_traffic(a.workload, dom, frac_samp)
    print(f"[bench] stage ms/step: " + ", ".join(f"{k} {v:.4f}" for k, v in prof.items()),
          file=sys.stderr, flush=True)
    by_kind = {}
    for w, pr, smp in profs:
        kind = "sampling" if smp else "plain"
        d_ms = pr[dom]
        by_kind[kind] = {"stage_ms": pr, "dominant_frac":
                         stage_bytes(dom, n, m, smp) / (d_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        print(f"[bench] {kind} steps: " + ", ".join(f"{k} {v:.4f}" for k, v in pr.items())
              + f"; {dom} at {by_kind[kind]['dominant_frac']:.3f} of HBM peak",
              file=sys.stderr, flush=True)

    cpu = None
    cpu_1core = None
    if not a.no_cpu_baseline and "batch" in wl:
        print("[bench] no cpu baseline for a batch of cases (the reference runs one case)",
              file=sys.stderr, flush=True)
    elif not a.no_cpu_baseline:
        cpu_steps = a.cpu_steps or cpu_sample_steps(wl, 2.4e8, 2000)
        cpu_1core = cpu_baseline(wl, cpu_steps, cpu_steps // 2)
        ranks = host_cores()
        mpi_wl = wl if mpi_fits(wl, ranks) else WORKLOADS["c2"]
        cpu = mpi_baseline(mpi_wl, ranks, share) or cpu_1core

    data = "synthetic: the freestream initial fill"
    if a.load_state:
        data += f", developed and saved earlier ({a.load_state})"
    elif developed:
        data += f", developed by this path's own steps 0..{developed - 1} before the warmup"
    out = {
        "metric": "particle-steps/s on 1 MI355X (vs CPU)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (positions, arithmetic and moment sums; velocities stored as f32)",
        "data": data,
        "config": {"workload": f"{a.workload}: {label}", "cases": nc,
                   "grid": [p.nx, p.ny] if nc == 1 else "per case", "cells": int(m),
                   "particles": int(np_end), "timed_steps": f"{step0}..{step0 + steps - 1}",
                   "sampling_after_step": first,
                   "sampling_steps_timed": n_samp,
                   "tuning": {"schedule": "small" if tuning[0] == 1 else "large",
                              "chunk": int(tuning[1]),
                              "particles_per_thread": int(tuning[2])}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src, "launch_ms": ms, "algorithmic_bytes": nbytes,
                     "kernel_bytes": kbytes,
                     "kernel_bytes_frac": kbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "step_roofline": {"bytes_per_step": sbytes, "achieved": sbytes / (wall / steps) / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": sbytes / (wall / steps) / 1e9 / HBM_PEAK_GBS,
                          "particles_per_step": psteps / steps, "cells": int(m),
                          "injected_per_step": inj_per_step, "collisions_per_step": coll_per_step},
        "stage_ms": prof,
        "stage_ms_by_kind": by_kind,
        "sampling_fraction_of_timed_steps": frac_samp,
        "cpu_baseline": cpu,
        "cpu_baseline_one_core": cpu_1core,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
