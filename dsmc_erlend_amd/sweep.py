"""Batched sweeps and the learned warm start, driven through the C-ABI.

A sweep is many independent cases (angle of attack, wall model, Mach, density, grid) run in one
particle pool (`Simulation([...])`, dsmc_create_batch). Each case draws the same random numbers
alone or in a batch, so a batched run is a faster way to get exactly the case-by-case results.

The warm start (include/dsmc_surrogate.h) is trained on steady fields the batched GPU DSMC
produces here, predicts (n, ux, uy, T) for new cases, and seeds their particles from those local
Maxwellians (dsmc_init_from_fields). The DSMC's sampled field stays the output, and it is what
the network's error is measured against: `field_error` compares any field with a converged
freestream-initialised run.

Command line: `python -m dsmc_erlend_amd.sweep train` makes the training fields and trains the
network (weights to a .npy file) and prints one JSON line. The measurement of held-out sweeps
(the network's field error, plain and warm-started sweeps timed to steady state) is
`sweep_bench.py` at the repository root, built on the functions here.

Steady state is decided on blocks of cells (`bins` x `bins`, default the 50 x 50 blocks of 0.8 m on
the 40 m domain) from the average over a trailing window of steps: per-cell sampling noise at
5-40 particles per cell would otherwise keep a cell-level 2% out of reach for either method, and a
running average from the first step would carry the initial transient for many thousand steps.
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


def bin_index(n, nb):
    """Block of each cell along one axis: floor((k + 0.5) nb / n), in integers."""
    return (2 * np.arange(n) + 1) * nb // (2 * n)


def cells_per_bin(p, bx, by):
    return (np.bincount(bin_index(p.nx, bx), minlength=bx)[:, None]
            * np.bincount(bin_index(p.ny, by), minlength=by)[None, :])


def moments_to_fields(s, nsteps, k, ncell):
    """(n, ux, uy, T) planes [4, bx, by] from CELL sums s [bx, by, 17] gathered over `nsteps`
    sampling steps in blocks of `ncell` cells: Printing's formulas (src/Sampling.cpp:152-196)
    with the block's volume. k: the case's dsmc_constants."""
    n0 = s[..., 0]
    ok = n0 > 0
    d = np.where(ok, n0, 1.0)
    f = np.zeros((4,) + n0.shape)
    f[0] = k.weight * n0 / (max(nsteps, 1) * k.dv * ncell)
    f[1] = np.where(ok, s[..., 1] / d, 0.0)
    f[2] = np.where(ok, s[..., 2] / d, 0.0)
    t = 0.0
    for a in range(3):
        u = s[..., 1 + a] / d
        t = t + k.molecule_mass * (s[..., 4 + 4 * a] / d - u * u) / KB
    f[3] = np.where(ok, t / 3.0, 0.0)
    return f


def reference_fields(cases, steps, first_sampling, bins, seed=None, log=None):
    """Converged freestream-initialised fields of a batch: per case the full-resolution planes
    [4, nx, ny] and the block fields [4, bins, bins] over the sampled steps."""
    cs = [p.replace(first_sampling_step=first_sampling, n_step=steps,
                    **({"seed": seed} if seed is not None else {})) for p in cases]
    with Simulation(cs) as sim:
        sim.initial_conditions()
        done = 0
        while done < steps:
            k = min(1000, steps - done)
            sim.run(k)
            done += k
            if log:
                log(f"  reference batch of {len(cs)}: step {done}/{steps}")
        full, binned = [], []
        for i, p in enumerate(cs):
            full.append(case_fields(sim, i))
            ss = sim.state(i).sample_step
            binned.append(moments_to_fields(sim.binned_cells(bins, bins, i), ss,
                                            sim.constants(i), cells_per_bin(p, bins, bins)))
        return full, binned


def run_to_steady(cases, refs, tol, window, check_every, max_steps, bins, init_fields=None,
                  seed=None, log=None):
    """Runs the cases as one batch, sampling every step, and every `check_every` steps compares
    each case's average over the trailing `window` steps (differences of its block sums) with
    its reference block field. Returns per case the first checked step at which
    max(e_n, e_u, e_T) < tol (or None), the GPU seconds up to that step, and the history."""
    cs = [p.replace(first_sampling_step=0, n_step=max_steps,
                    **({"seed": seed} if seed is not None else {})) for p in cases]
    ncell = [cells_per_bin(p, bins, bins) for p in cs]
    hit_step = [None] * len(cs)
    hit_time = [None] * len(cs)
    hist = [[] for _ in cs]
    with Simulation(cs) as sim:
        kc = [sim.constants(i) for i in range(len(cs))]
        t0 = time.perf_counter()
        if init_fields is None:
            sim.initial_conditions()
        else:
            sim.init_from_fields(init_fields)
        gpu = time.perf_counter() - t0
        snaps = [[(0, np.zeros((bins, bins, 17)), 0)] for _ in cs]
        step = 0
        while step < max_steps and any(h is None for h in hit_step):
            k = min(check_every, max_steps - step)
            t0 = time.perf_counter()
            sim.run(k)
            gpu += time.perf_counter() - t0
            step += k
            for i in range(len(cs)):
                if hit_step[i] is not None:
                    continue
                snaps[i].append((step, sim.binned_cells(bins, bins, i), sim.state(i).sample_step))
                while snaps[i][0][0] < step - window:
                    snaps[i].pop(0)
                s0, S0, n0 = snaps[i][0]
                if step - s0 < window:
                    continue
                f = moments_to_fields(snaps[i][-1][1] - S0, snaps[i][-1][2] - n0, kc[i], ncell[i])
                e = field_error(f, refs[i])
                hist[i].append((step, e))
                if max(e) < tol:
                    hit_step[i], hit_time[i] = step, gpu
            if log and step % (10 * check_every) == 0:
                log(f"  {sum(h is not None for h in hit_step)}/{len(cs)} steady at step {step}")
    return hit_step, hit_time, hist


# ---------------------------------------------------------------------------------------
def _log(msg):
    print(f"[sweep] {msg}", file=sys.stderr, flush=True)


def train_experiment(a):
    """Training fields from the batched GPU DSMC on the network raster, then the network."""
    from .surrogate import Surrogate
    out = {"experiment": "warm-start training", "config": vars(a).copy()}
    R = a.raster
    t0 = time.perf_counter()
    train = draw_cases(a.train_cases, a.seed, grid=R, npc=a.train_npc, seed_base=100)
    ftrain, rates = [], []
    for b0 in range(0, len(train), a.batch_cases):
        f, rate = steady_fields(train[b0:b0 + a.batch_cases], a.train_steps,
                                a.train_steps - a.train_sample)
        ftrain += f
        rates.append(rate)
        _log(f"training fields {b0 + len(f)}/{len(train)} ({rate:.3e} particle-steps/s)")
    out["train_data_s"] = time.perf_counter() - t0
    out["train_data_rate"] = float(np.mean(rates))
    ftrain = np.stack(ftrain).astype(np.float32)
    nval = max(1, len(train) // 16)
    with Surrogate(height=R, width=R, max_batch=a.net_batch, seed=a.seed + 1) as net:
        net.set_dataset(train[nval:], ftrain[nval:])
        val0 = net.evaluate(train[:nval], ftrain[:nval])
        hist = net.train(epochs=a.epochs, batch=a.net_batch, lr=a.lr, seed=a.seed + 2)
        ms, flops = net.last_train_stats()
        val1 = net.evaluate(train[:nval], ftrain[:nval])
        np.save(a.weights, net.weights())
        out["train"] = {"samples": len(train) - nval, "steps": int(hist.size), "ms": ms,
                        "conv_tflops": flops / (ms * 1e-3) / 1e12,
                        "loss_first": float(hist[:10].mean()),
                        "loss_last": float(hist[-10:].mean()),
                        "val_loss_init": val0, "val_loss": val1, "params": net.num_params()}
    _log(f"trained {hist.size} steps in {ms / 1e3:.1f} s ({out['train']['conv_tflops']:.1f} "
         f"TFLOP/s), loss {out['train']['loss_first']:.4f} -> {out['train']['loss_last']:.4f}, "
         f"validation {val0:.4f} -> {val1:.4f}")
    print(json.dumps(out), flush=True)


def snapshot_run(p, steps, every, bins, init_fields=None):
    """One case run by itself from the freestream fill (or from `init_fields`, [4, H, W]),
    sampling every step: the block sums every `every` steps (index 0 = before the first step),
    the GPU seconds of its steps and of its fill."""
    p = p.replace(first_sampling_step=0, n_step=steps)
    with Simulation([p]) as sim:
        t0 = time.perf_counter()
        if init_fields is None:
            sim.initial_conditions()
        else:
            sim.init_from_fields(init_fields[None])
        t_init = time.perf_counter() - t0
        snaps = [(0, sim.binned_cells(bins, bins, 0), sim.state(0).sample_step)]
        t_steps = 0.0
        done = 0
        while done < steps:
            k = min(every, steps - done)
            t0 = time.perf_counter()
            sim.run(k)
            t_steps += time.perf_counter() - t0
            done += k
            snaps.append((done, sim.binned_cells(bins, bins, 0), sim.state(0).sample_step))
        return snaps, t_steps / steps, t_init, sim.constants(0)


def window_field(snaps, a, b, k, ncell):
    """Field averaged over the steps between snapshots a and b."""
    return moments_to_fields(snaps[b][1] - snaps[a][1], snaps[b][2] - snaps[a][2], k, ncell)


class SteadyTest:
    """The steady test of one running case from its block sums at every check, keeping only
    the snapshots the rule can still use: under "second_half" those from index b//2 on (b//2
    never decreases), under "hindsight" the last `max_starts` (sampling starts more than that
    many checks back are not tried)."""

    def __init__(self, ref, k, ncell, rule, max_starts=40):
        self.ref, self.k, self.ncell, self.rule, self.max_starts = ref, k, ncell, rule, max_starts
        self.snaps = {}    # snapshot index -> (block sums [bx, by, 17], sampled steps)
        self.b = -1

    def add(self, sums, sampled):
        """Adds the snapshot of the next check; returns the best error (worst channel) over
        the windows [F, N] the rule allows, or None before there is a window."""
        self.b += 1
        self.snaps[self.b] = (sums, sampled)
        lo = self.b // 2 if self.rule == "second_half" else self.b - self.max_starts
        for a in [a for a in self.snaps if a < lo]:
            del self.snaps[a]
        if self.b == 0:
            return None
        starts = [self.b // 2] if self.rule == "second_half" else \
            [a for a in sorted(self.snaps) if a < self.b]
        last, n_last = self.snaps[self.b]
        fs = [moments_to_fields(last - self.snaps[a][0], n_last - self.snaps[a][1], self.k,
                                self.ncell) for a in starts]
        return min(max(field_error(f, self.ref)) for f in fs)


def cell_moments(cells, nsteps, k):
    """Per cell (n, ux, uy, Tx, Ty, Tz) [M, 6] from CELL sums [M, 17] over `nsteps` sampled
    steps (Printing's formulas, src/Sampling.cpp:152-196, with a temperature per component);
    0 where a cell saw no particle."""
    n0 = cells[:, 0]
    ok = n0 > 0
    d = np.where(ok, n0, 1.0)
    out = np.zeros((cells.shape[0], 6))
    out[:, 0] = k.weight * n0 / (max(nsteps, 1) * k.dv)
    for a in range(3):
        u = cells[:, 1 + a] / d
        if a < 2:
            out[:, 1 + a] = np.where(ok, u, 0.0)
        out[:, 3 + a] = np.where(ok, np.maximum(
            k.molecule_mass * (cells[:, 4 + 4 * a] / d - u * u) / KB, 0.0), 0.0)
    return out


def retiring_batch(cases, refs, tol, rule, check_every, max_steps, bins, init_fields=None,
                   init_moments=None, log=None):
    """Runs the cases as one batch, sampling every step, and retires each case
    (dsmc_retire_case) at the first check at which it is steady by `rule`, so the pool shrinks
    to the cases still running. The pool starts from the freestream fill, from `init_fields`
    ([ncases, 4, w, h] rasters of n, ux, uy, T) or from `init_moments` ([cells of all cases,
    6]: n, ux, uy, Tx, Ty, Tz per cell). Returns per case the step it retired at (None: never,
    it ran to `max_steps`), the fill's and the steps' GPU seconds, and per check interval the
    cases running and its seconds per step."""
    cs = [p.replace(first_sampling_step=0, n_step=max_steps) for p in cases]
    ncell = [cells_per_bin(p, bins, bins) for p in cs]
    done_at = [None] * len(cs)
    intervals = []
    with Simulation(cs) as sim:
        kc = [sim.constants(i) for i in range(len(cs))]
        t0 = time.perf_counter()
        if init_moments is not None:
            sim.init_from_cell_moments(init_moments)
        elif init_fields is not None:
            sim.init_from_fields(init_fields)
        else:
            sim.initial_conditions()
        t_init = time.perf_counter() - t0
        tests = [SteadyTest(refs[i], kc[i], ncell[i], rule) for i in range(len(cs))]
        for i, t in enumerate(tests):
            t.add(sim.binned_cells(bins, bins, i), sim.state(i).sample_step)
        active = list(range(len(cs)))
        step, gpu = 0, 0.0
        while step < max_steps and active:
            k = min(check_every, max_steps - step)
            t0 = time.perf_counter()
            sim.run(k)
            dt = time.perf_counter() - t0
            gpu += dt
            step += k
            intervals.append((step, len(active), dt / k))
            for i in list(active):
                e = tests[i].add(sim.binned_cells(bins, bins, i), sim.state(i).sample_step)
                if e is not None and e < tol:
                    sim.retire(i)
                    done_at[i] = step
                    active.remove(i)
                    tests[i] = None
            if log and step % (10 * check_every) == 0:
                log(f"  step {step}: {len(active)} of {len(cs)} cases running, "
                    f"{1e3 * dt / k:.2f} ms per step")
    return done_at, t_init, gpu, intervals


def steady_steps(snaps, ref, k, ncell, tols):
    """For each tolerance: the first snapshot step N at which some sampling start F < N (on the
    snapshot grid: FirstSamplingStep chosen in hindsight) gives an average over [F, N] within
    the tolerance of `ref` (worst of e_n, e_u, e_T); and the same with the usual rule of
    sampling the second half, F = N/2. None where no window gets there. Also the best error
    reached at each N."""
    best = []
    half = []
    for b in range(1, len(snaps)):
        errs = [max(field_error(window_field(snaps, a, b, k, ncell), ref)) for a in range(b)]
        best.append((snaps[b][0], min(errs)))
        half.append((snaps[b][0], errs[b // 2]))
    out = {}
    for tol in tols:
        out[f"tol_{tol}"] = {
            "hindsight": next((n for n, e in best if e < tol), None),
            "second_half": next((n for n, e in half if e < tol), None)}
    return out, best, half


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m dsmc_erlend_amd.sweep")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="training fields from the batched GPU DSMC, then the network")
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--raster", type=int, default=256)
    t.add_argument("--train-cases", type=int, default=512)
    t.add_argument("--train-npc", type=float, default=5.0)
    t.add_argument("--train-steps", type=int, default=12000)
    t.add_argument("--train-sample", type=int, default=4000)
    t.add_argument("--batch-cases", type=int, default=128)
    t.add_argument("--net-batch", type=int, default=8)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--weights", default="warmstart_weights.npy")
    a = ap.parse_args(argv)
    train_experiment(a)


if __name__ == "__main__":
    main()
