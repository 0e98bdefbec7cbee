"""Benchmark: particle-steps/s of the DSMC step on one MI355X.

Workload (default): the reference's default case on one GPU - 100x100 grid on [-20,20]^2 m,
4x4 m box, NPCFree 10 (~1e5 particles), Ma 4, n 1e17 m^-3, T 200 K, AoA 15 deg, diffuse walls
at 300 K, Dt 1e-5 s, sampling every step after step 10000 of 20000. Synthetic data: the
freestream initial fill. W warmup steps, then K timed steps of the fused, graph-captured
step; value = sum over the timed steps of the particles indexed that step / wall-clock.

After the timed region: an eager pass with HIP events around each stage gives the dominant
kernel's mean duration for the roofline; the CPU baseline times the reference's own step
functions (oracle/_ref, compiled from its sources) or, where that binary cannot run, the
oracle's restatement of them, on a bounded sample of the same case.
"""
import argparse
import json
import os
import subprocess
import sys
import time

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
}


def stage_bytes(stage, n, m, sampling):
    """Algorithmic HBM bytes of one launch of a stage's kernel: every array element read or
    written once (FP64 x,y and FP32 vx,vy,vz; u32 keys, offsets, positions)."""
    if stage == "move":      # R x,y 16 + vx,vy 8; W x,y 16; W key 4; counts R/W 8 per cell
        return 44 * n + 8 * m
    if stage == "scatter":   # R key 4, W perm 4 per particle; cursors R/W 8 per cell
        return 8 * n + 8 * m
    if stage == "gather":    # R perm 4 + R 28 + W 28 per particle; offsets and (sigma*g)max;
        # on sampling steps 13 FP64 sums read and written per cell
        return 60 * n + 16 * m + (208 * m if sampling else 0)
    if stage == "scan":      # counts R, offsets and cursors W, counts cleared
        return 16 * m
    return 0


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
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only a copy kernel of known size (for PMC byte counters)")
    a = ap.parse_args()
    if a.calibrate:
        calibrate()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the path shards "
              "by whole cases, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from dsmc_erlend_amd import Simulation, default_params, STAGES
    spec = dict(WORKLOADS[a.workload])
    label = spec.pop("label")
    total = spec["n_step"]
    warmup = a.warmup if a.warmup is not None else 100
    steps = a.steps if a.steps is not None else max(1, total - warmup)
    p = default_params(**spec)
    sim = Simulation(p)
    t_init = time.perf_counter()
    sim.initial_conditions()
    t_init = time.perf_counter() - t_init
    print(f"[bench] {a.workload}: {sim.num_particles()} particles after init "
          f"({t_init:.2f} s)", file=sys.stderr, flush=True)
    if warmup:
        sim.run(warmup)
    s0 = sim.state()
    t0 = time.perf_counter()
    sim.run(steps)              # synchronises at the end
    t1 = time.perf_counter()
    s1 = sim.state()
    psteps = s1.particle_steps - s0.particle_steps
    wall = t1 - t0
    value = psteps / wall
    print(f"[bench] timed {steps} steps in {wall:.3f} s: {value:.4e} particle-steps/s, "
          f"{s1.np} particles", file=sys.stderr, flush=True)

    # dominant kernel: eager steps with events around each stage (after the timed region)
    prof = sim.profile(a.profile_steps)
    n = sim.num_particles()
    m = p.nx * p.ny
    sampling = sim.state().step > p.first_sampling_step
    # single-kernel stages (the scan is three launches on large grids)
    kernels = {k: v for k, v in prof.items() if k in ("move", "gather", "scatter")}
    dom = max(kernels, key=kernels.get)
    ms = prof[dom]
    nbytes = stage_bytes(dom, n, m, sampling)
    achieved = nbytes / (ms * 1e-3) / 1e9
    print(f"[bench] stage ms/step: " + ", ".join(f"{k} {v:.4f}" for k, v in prof.items()),
          file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(WORKLOADS[a.workload], a.cpu_steps, a.cpu_steps // 2)

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
        "dtype": "f64",
        "data": "synthetic (freestream initial fill)",
        "config": {"workload": f"{a.workload}: {label}", "grid": [p.nx, p.ny],
                   "particles": int(s1.np), "timed_steps": f"{warmup}..{warmup + steps - 1}",
                   "sampling_after_step": p.first_sampling_step},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "launch_ms": ms, "algorithmic_bytes": nbytes},
        "stage_ms": prof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
