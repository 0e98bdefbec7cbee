"""Batched sweeps and the learned warm start, driven through the C-ABI.

A sweep is many independent cases (angle of attack, wall model, Mach, density, grid) run in one
particle pool (`Simulation([...])`, dsmc_create_batch). Each case draws the same random numbers
alone or in a batch, so a batched run is a faster way to get exactly the case-by-case results.

The warm start (include/dsmc_surrogate.h) is trained on steady fields the batched GPU DSMC
produces here, predicts (n, ux, uy, T) for new cases, and seeds their particles from those local
Maxwellians (dsmc_init_from_fields). The DSMC's sampled field stays the output and the judge:
`field_error` measures any field against a converged freestream-initialised run.

Command line: `python -m dsmc_erlend_amd.sweep warmstart [options]` runs the whole experiment
(training data, training, held-out evaluation, time to steady state) and prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import Simulation, default_params, WALL_CLL, WALL_DIFFUSE, WALL_SPECULAR

KB = 1.380662E-23
NA = 6.022045E+23


def freestream(p):
    m = p.molar_mass / NA
    return p.n_free, p.mach * np.sqrt(5.0 / 3.0 * KB * p.t_free / m), p.t_free


def draw_cases(n, seed, grid=None, npc=None, seed_base=1000):
    """Cases of the batched sweep: AoA uniform in [-30, 30] deg; wall model cycled specular,
    diffuse, CLL with acc_n, acc_t uniform in [0.5, 1]; grid side uniform in [200, 800] (or
    `grid`); NPCFree uniform in [5, 40] (or `npc`); Mach uniform in [2, 8]; n log-uniform in
    [1e16, 1e18]. Every case gets its own case_id, so its random streams are its own."""
    rng = np.random.default_rng(seed)
    walls = [WALL_SPECULAR, WALL_DIFFUSE, WALL_CLL]
    out = []
    for i in range(n):
        side = int(grid) if grid else int(rng.integers(200, 801))
        p = default_params(
            nx=side, ny=side,
            npc_free=float(npc) if npc else float(rng.uniform(5, 40)),
            attack_angle_deg=float(rng.uniform(-30, 30)),
            wall_model=walls[i % 3],
            acc_n=float(rng.uniform(0.5, 1.0)), acc_t=float(rng.uniform(0.5, 1.0)),
            mach=float(rng.uniform(2, 8)),
            n_free=float(10 ** rng.uniform(16, 18)),
            seed=seed_base + seed, case_id=i,
            capacity_factor=2.5)
        out.append(p)
    return out


def case_fields(sim, i):
    """Running-average (n, ux, uy, T) of case i as planes [4, nx, ny] (x-major)."""
    p = sim.params[i]
    f = sim.fields(i).reshape(p.nx, p.ny, 6)
    return np.stack([f[..., 0], f[..., 1], f[..., 2], f[..., 5]])


def coarsen(f, side):
    """Block average of planes [4, nx, ny] onto side x side bins (a cell goes to the bin holding
    its centre); velocity and temperature weighted by density, as the moments they come from."""
    _, nx, ny = f.shape
    bx = ((np.arange(nx) + 0.5) * side / nx).astype(int)
    by = ((np.arange(ny) + 0.5) * side / ny).astype(int)
    idx = (bx[:, None] * side + by[None, :]).ravel()
    n = f[0].ravel()
    cnt = np.bincount(idx, minlength=side * side)
    nsum = np.bincount(idx, weights=n, minlength=side * side)
    out = np.zeros((4, side * side))
    out[0] = nsum / np.maximum(cnt, 1)
    for q in (1, 2, 3):
        s = np.bincount(idx, weights=n * f[q].ravel(), minlength=side * side)
        out[q] = np.where(nsum > 0, s / np.where(nsum > 0, nsum, 1), 0.0)
    return out.reshape(4, side, side)


def field_error(f, ref):
    """Relative L2 errors of density, velocity (the vector) and temperature over gas cells
    (n_ref > 0): (e_n, e_u, e_T)."""
    gas = ref[0] > 0
    if not gas.any():
        return (np.inf, np.inf, np.inf)

    def rel(a, b):
        d = np.sqrt(np.sum((a - b) ** 2))
        return d / max(np.sqrt(np.sum(b ** 2)), 1e-300)

    e_n = rel(f[0][gas], ref[0][gas])
    e_u = rel(np.stack([f[1][gas], f[2][gas]]), np.stack([ref[1][gas], ref[2][gas]]))
    e_t = rel(f[3][gas], ref[3][gas])
    return (float(e_n), float(e_u), float(e_t))


def resample(f, H, W):
    """Bilinear resample of planes [4, nx, ny] onto an H x W raster of the same domain, by cell
    centres (the mapping dsmc_init_from_fields uses in the other direction)."""
    _, nx, ny = f.shape
    fx = (np.arange(H) + 0.5) / H * nx - 0.5
    fy = (np.arange(W) + 0.5) / W * ny - 0.5
    a0 = np.clip(np.floor(fx).astype(int), 0, nx - 1)
    b0 = np.clip(np.floor(fy).astype(int), 0, ny - 1)
    ta = np.clip(fx - a0, 0.0, 1.0)
    tb = np.clip(fy - b0, 0.0, 1.0)
    a1 = np.minimum(a0 + 1, nx - 1)
    b1 = np.minimum(b0 + 1, ny - 1)
    ta, tb = ta[:, None], tb[None, :]
    return ((1 - ta) * ((1 - tb) * f[:, a0][:, :, b0] + tb * f[:, a0][:, :, b1])
            + ta * ((1 - tb) * f[:, a1][:, :, b0] + tb * f[:, a1][:, :, b1]))


def freestream_fields(p, nx=None, ny=None):
    """The freestream fill as a field (the method without the network)."""
    nx, ny = nx or p.nx, ny or p.ny
    nf, uf, tf = freestream(p)
    a = np.radians(p.attack_angle_deg)
    f = np.zeros((4, nx, ny))
    f[0], f[1], f[2], f[3] = nf, uf * np.cos(a), -uf * np.sin(a), tf
    return f


def steady_fields(cases, steps, first_sampling, init_fields=None, seed=None, log=None):
    """Runs the cases as one batch for `steps` steps (sampling after `first_sampling`) from the
    freestream fill or from `init_fields` [ncases, 4, w, h]; returns their fields and the
    particle-steps/s of the run."""
    cs = [p.replace(first_sampling_step=first_sampling, n_step=steps,
                    **({"seed": seed} if seed is not None else {})) for p in cases]
    with Simulation(cs) as sim:
        if init_fields is None:
            sim.initial_conditions()
        else:
            sim.init_from_fields(init_fields)
        s0 = sum(sim.state(i).particle_steps for i in range(len(cs)))
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            k = min(1000, steps - done)
            sim.run(k)
            done += k
            if log:
                log(f"  batch of {len(cs)}: step {done}/{steps}")
        wall = time.perf_counter() - t0
        ps = sum(sim.state(i).particle_steps for i in range(len(cs))) - s0
        return [case_fields(sim, i) for i in range(len(cs))], ps / wall


def run_to_steady(cases, refs, tol, first_sampling, check_every, max_steps, init_fields=None,
                  side=None, seed=None):
    """Runs the cases as one batch and checks every `check_every` steps how far each case's
    running average is from its reference field (`side`: compared on side x side bins). Returns
    per case the first checked step at which max(e_n, e_u, e_T) < tol (or None), the GPU
    seconds spent up to that step, and the error history."""
    cs = [p.replace(first_sampling_step=first_sampling, n_step=max_steps,
                    **({"seed": seed} if seed is not None else {})) for p in cases]
    rc = [coarsen(r, side) if side else r for r in refs]
    hit_step = [None] * len(cs)
    hit_time = [None] * len(cs)
    hist = [[] for _ in cs]
    with Simulation(cs) as sim:
        t0 = time.perf_counter()
        if init_fields is None:
            sim.initial_conditions()
        else:
            sim.init_from_fields(init_fields)
        gpu = time.perf_counter() - t0
        step = 0
        while step < max_steps and any(h is None for h in hit_step):
            k = min(check_every, max_steps - step)
            t0 = time.perf_counter()
            sim.run(k)
            gpu += time.perf_counter() - t0
            step += k
            if step <= first_sampling:
                continue
            for i in range(len(cs)):
                if hit_step[i] is not None:
                    continue
                f = case_fields(sim, i)
                e = field_error(coarsen(f, side) if side else f, rc[i])
                hist[i].append((step, e))
                if max(e) < tol:
                    hit_step[i], hit_time[i] = step, gpu
    return hit_step, hit_time, hist


# ---------------------------------------------------------------------------------------
def _log(msg):
    print(f"[sweep] {msg}", file=sys.stderr, flush=True)


def warmstart_experiment(a):
    from .surrogate import Surrogate
    out = {"experiment": "learned warm start", "config": vars(a).copy()}
    R = a.raster
    # 1. training fields from the batched GPU DSMC, on the network raster
    t0 = time.perf_counter()
    train = draw_cases(a.train_cases, a.seed, grid=R, npc=a.train_npc, seed_base=100)
    ftrain, rates = [], []
    for b0 in range(0, len(train), a.batch_cases):
        chunk = train[b0:b0 + a.batch_cases]
        f, rate = steady_fields(chunk, a.train_steps, a.train_steps - a.train_sample)
        ftrain += f
        rates.append(rate)
        _log(f"training fields {b0 + len(chunk)}/{len(train)} ({rate:.3e} particle-steps/s)")
    out["train_data_s"] = time.perf_counter() - t0
    out["train_data_rate"] = float(np.mean(rates))
    ftrain = np.stack(ftrain).astype(np.float32)
    # 2. the network, trained on the GPU by the library's own kernels
    net = Surrogate(height=R, width=R, max_batch=a.net_batch, seed=a.seed + 1)
    net.set_dataset(train, ftrain)
    hist = net.train(epochs=a.epochs, batch=a.net_batch, lr=a.lr, seed=a.seed + 2)
    ms, flops = net.last_train_stats()
    out["train"] = {"steps": int(hist.size), "ms": ms, "tflops": flops / (ms * 1e-3) / 1e12,
                    "loss_first": float(hist[:10].mean()), "loss_last": float(hist[-10:].mean()),
                    "params": net.num_params()}
    _log(f"trained {hist.size} steps in {ms / 1e3:.1f} s ({out['train']['tflops']:.1f} TFLOP/s),"
         f" loss {out['train']['loss_first']:.4f} -> {out['train']['loss_last']:.4f}")
    # 3. held-out sweep cases: converged plain reference fields (their own seed)
    test = draw_cases(a.test_cases, a.seed + 7, seed_base=200)
    t0 = time.perf_counter()
    refs, _ = steady_fields(test, a.ref_steps, a.ref_steps - a.ref_sample, seed=9999)
    out["ref_s"] = time.perf_counter() - t0
    _log(f"reference fields of {len(test)} held-out cases in {out['ref_s']:.1f} s")
    t0 = time.perf_counter()
    pred = net.predict(test)
    out["predict_s"] = time.perf_counter() - t0
    # network error against the converged DSMC field, on the raster, next to the freestream's
    err_net, err_free = [], []
    for i, p in enumerate(test):
        r = resample(refs[i], R, R)
        err_net.append(field_error(pred[i], r))
        err_free.append(field_error(freestream_fields(p, R, R), r))
    out["net_error"] = np.mean(err_net, axis=0).tolist()
    out["freestream_error"] = np.mean(err_free, axis=0).tolist()
    _log(f"held-out field error (n, u, T): network {out['net_error']}, "
         f"freestream {out['freestream_error']}")
    # 4. time to steady: plain (freestream, sampling after a fixed transient) case by case,
    # against the batched warm start (sampling after a short relaxation)
    side = a.compare_side or None
    plain_steps, plain_time = [], []
    if not a.skip_case_by_case:
        for i, p in enumerate(test):
            s, tm, _ = run_to_steady([p], [refs[i]], a.tol, a.plain_first, a.check_every,
                                     a.max_steps, side=side)
            plain_steps.append(s[0])
            plain_time.append(tm[0])
            _log(f"plain case {i}: steady at step {s[0]} ({tm[0]})")
    ws, wt, _ = run_to_steady(test, refs, a.tol, a.warm_first, a.check_every, a.max_steps,
                              init_fields=pred, side=side)
    bs, bt, _ = run_to_steady(test, refs, a.tol, a.plain_first, a.check_every, a.max_steps,
                              side=side)
    out["steady"] = {
        "tol": a.tol, "compare_side": side,
        "plain_case_by_case_steps": plain_steps, "plain_case_by_case_s": plain_time,
        "warm_batched_steps": ws, "warm_batched_s": wt[-1] if all(wt) else None,
        "warm_batched_case_s": wt,
        "plain_batched_steps": bs, "plain_batched_s": bt[-1] if all(bt) else None,
    }
    if plain_time and all(t is not None for t in plain_time) and all(wt):
        total_plain = float(np.sum(plain_time))
        total_warm = max(wt) + out["predict_s"]
        out["steady"]["speedup"] = total_plain / total_warm
    print(json.dumps(out), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m dsmc_erlend_amd.sweep")
    sub = ap.add_subparsers(dest="cmd", required=True)
    w = sub.add_parser("warmstart", help="train the warm start and measure its gain")
    w.add_argument("--seed", type=int, default=0)
    w.add_argument("--raster", type=int, default=256)
    w.add_argument("--train-cases", type=int, default=512)
    w.add_argument("--train-npc", type=float, default=10.0)
    w.add_argument("--train-steps", type=int, default=6000)
    w.add_argument("--train-sample", type=int, default=2000)
    w.add_argument("--batch-cases", type=int, default=64)
    w.add_argument("--net-batch", type=int, default=8)
    w.add_argument("--epochs", type=int, default=40)
    w.add_argument("--lr", type=float, default=1e-3)
    w.add_argument("--test-cases", type=int, default=64)
    w.add_argument("--ref-steps", type=int, default=12000)
    w.add_argument("--ref-sample", type=int, default=8000)
    w.add_argument("--tol", type=float, default=0.02)
    w.add_argument("--plain-first", type=int, default=2000)
    w.add_argument("--warm-first", type=int, default=200)
    w.add_argument("--check-every", type=int, default=250)
    w.add_argument("--max-steps", type=int, default=12000)
    w.add_argument("--compare-side", type=int, default=100)
    w.add_argument("--skip-case-by-case", action="store_true")
    a = ap.parse_args(argv)
    if a.cmd == "warmstart":
        warmstart_experiment(a)


if __name__ == "__main__":
    main()
