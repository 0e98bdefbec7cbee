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
network (weights to a .npy file); `python -m dsmc_erlend_amd.sweep evaluate` runs held-out sweep
cases to converged fields, measures the network's field error, and times plain and warm-started
sweeps to steady state. Each prints one JSON line.

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


def _save_json(path, obj):
    with open(path, "w") as f:
        json.dump(obj, f)


def parse_cases(spec, n):
    """Case indices from "0-15", "3,7,9" or None (all n)."""
    if not spec:
        return list(range(n))
    out = []
    for part in spec.split(","):
        if "-" in part:
            lo, hi = part.split("-")
            out.extend(range(int(lo), int(hi) + 1))
        else:
            out.append(int(part))
    return [i for i in out if 0 <= i < n]


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


def evaluate_experiment(a):
    """Held-out sweep cases, in phases that keep their results in `--state` (each phase fits
    one bounded GPU session): `ref` runs the converged plain fields and, with weights, the
    network's field error against them.

    `cases` (for the cases in `--cases`) and `summary` are the case-by-case protocol: each case
    runs by itself, first a long freestream reference (`--ref-steps`, its own seed, averaged
    over the last `--ref-sample` steps; the two halves of that window are compared as a check
    of convergence), then from the freestream fill and from the network's fields up to
    `--max-steps`, with block sums every `--check-every` steps. A run is steady at the first
    step N at which an average over [F, N] is within the tolerance of the reference, with F
    either chosen in hindsight on the same grid or F = N/2; its wall-clock is its fill plus N
    times its own measured step time. `summary` adds these up over the cases found.

    The batched form: `batchref` makes the same references for the cases in `--cases` as one
    batch; `batched` runs all cases in one pool from the network's fields (`--kind warm`) or the
    freestream fill and retires each once it is steady by `--rule` at `--tol`; `replay` times
    the plain case-by-case runs with the current build for the steps they needed; `retime`
    measures each case's step time alone on a developed flow. `summary` then sets the batched
    runs against the plain runs case by case."""
    import os
    from .surrogate import Surrogate
    os.makedirs(a.state, exist_ok=True)
    R, B = a.raster, a.bins
    test = draw_cases(a.test_cases, a.seed + 7, seed_base=200)
    out = {"experiment": f"warm-start sweep, phase {a.phase}", "config": vars(a).copy(),
           "cases": len(test), "particles": int(sum(p.nx * p.ny * p.npc_free for p in test)),
           "cells": int(sum(p.nx * p.ny for p in test))}
    refs_path = os.path.join(a.state, "refs.npz")

    def predict():
        with Surrogate(height=R, width=R, max_batch=8, seed=0) as net:
            net.set_weights(np.load(a.weights))
            t0 = time.perf_counter()
            pred = net.predict(test)
            return pred, time.perf_counter() - t0

    if a.phase == "ref":
        t0 = time.perf_counter()
        full, refs = reference_fields(test, a.ref_steps, a.ref_steps - a.ref_sample, B,
                                      seed=9999, log=_log)
        out["ref_s"] = time.perf_counter() - t0
        np.savez(refs_path, refs=np.stack(refs))
        if a.weights and os.path.exists(a.weights):
            pred, out["predict_s"] = predict()
            err_net = [field_error(pred[i], resample(full[i], R, R)) for i in range(len(test))]
            err_free = [field_error(freestream_fields(p, R, R), resample(full[i], R, R))
                        for i, p in enumerate(test)]
            out["net_error"] = np.mean(err_net, axis=0).tolist()
            out["freestream_error"] = np.mean(err_free, axis=0).tolist()
            out["net_error_cases"] = [[round(x, 4) for x in e] for e in err_net]
            out["freestream_error_cases"] = [[round(x, 4) for x in e] for e in err_free]
            _log(f"field error (n, u, T): network {out['net_error']}, "
                 f"freestream {out['freestream_error']}")
    elif a.phase == "cases":
        sel = parse_cases(a.cases, len(test))
        pred = None
        if a.weights and os.path.exists(a.weights):
            with Surrogate(height=R, width=R, max_batch=8, seed=0) as net:
                net.set_weights(np.load(a.weights))
                pred = net.predict([test[i] for i in sel])
        tols = (0.02, 0.03, 0.05)
        for j, i in enumerate(sel):
            p = test[i]
            nc = cells_per_bin(p, B, B)
            t0 = time.perf_counter()
            case_path = os.path.join(a.state, f"case_{i}.json")
            ref_path = os.path.join(a.state, f"ref_{i}.npy")
            if a.warm_only and os.path.exists(ref_path):
                ref = np.load(ref_path)
                with Simulation([p.replace(n_step=1)]) as sim:
                    k = sim.constants(0)
                drift = None
            else:
                # the reference run is deterministic: a rerun reproduces the saved one
                ref_snaps, _, _, k = snapshot_run(p.replace(seed=9999), a.ref_steps,
                                                  a.ref_sample // 2, B)
                n = len(ref_snaps) - 1
                ref = window_field(ref_snaps, n - 2, n, k, nc)
                np.save(ref_path, ref)
                drift = field_error(window_field(ref_snaps, n - 1, n, k, nc),
                                    window_field(ref_snaps, n - 2, n - 1, k, nc))
            if a.warm_only:
                # the freestream run of an earlier `cases` run
                case = json.load(open(case_path))
            else:
                case = {"case": i, "nx": p.nx, "npc_free": p.npc_free, "mach": p.mach,
                        "n_free": p.n_free, "aoa": p.attack_angle_deg, "wall": p.wall_model,
                        "ref_steps": a.ref_steps, "ref_sample": a.ref_sample,
                        "ref_halves_error": [round(x, 5) for x in drift]}
            for kind in ("plain", "warm"):
                if kind == "plain" and a.warm_only:
                    continue
                if kind == "warm" and pred is None:
                    continue
                snaps, step_s, init_s, _ = snapshot_run(
                    p, a.max_steps, a.check_every, B,
                    init_fields=None if kind == "plain" else pred[j])
                steady, best, half = steady_steps(snaps, ref, k, nc, tols)
                case[kind] = {"step_s": step_s, "init_s": init_s, "steady": steady,
                              "best_error": [(s_, round(e, 5)) for s_, e in best],
                              "second_half_error": [(s_, round(e, 5)) for s_, e in half]}
            _save_json(case_path, case)
            _log(f"case {i}: {time.perf_counter() - t0:.1f} s, reference halves "
                 f"{case['ref_halves_error']}, plain {case['plain']['steady']}, "
                 f"warm {case.get('warm', {}).get('steady')}")
        out["cases_done"] = sel
    elif a.phase == "seeding":
        # how much of the warm start's transient the network causes: the selected cases
        # seeded from the network's fields, from their own converged field (through the
        # network's raster and cell by cell) and from the freestream fill; errors of
        # consecutive `--check-every`-step windows against the converged field
        sel = parse_cases(a.cases, len(test))
        pred = None
        if a.weights and os.path.exists(a.weights):
            with Surrogate(height=R, width=R, max_batch=8, seed=0) as net:
                net.set_weights(np.load(a.weights))
                pred = net.predict([test[i] for i in sel])
        res = {}
        for j, i in enumerate(sel):
            p = test[i]
            nc = cells_per_bin(p, B, B)
            with Simulation([p.replace(seed=9999, first_sampling_step=a.ref_steps - a.ref_sample,
                                       n_step=a.ref_steps)]) as sim:
                sim.initial_conditions()
                for done in range(0, a.ref_steps, 1000):
                    sim.run(min(1000, a.ref_steps - done))
                full = np.nan_to_num(case_fields(sim, 0))
                # per-component temperatures: T = (Tx + Ty + Tz) / 3 (src/Sampling.cpp:152-196)
                f6 = np.nan_to_num(sim.fields(0).reshape(-1, 6))
                tz = 3.0 * f6[:, 5] - f6[:, 3] - f6[:, 4]
                temps = np.stack([f6[:, 3], f6[:, 4], tz], axis=1)
                temps = np.where(temps > 0.0, temps, f6[:, 5:6])
                moments = np.concatenate([f6[:, :3], temps], axis=1)
                k = sim.constants(0)
                ref = moments_to_fields(sim.binned_cells(B, B, 0), sim.state(0).sample_step, k,
                                        nc)
            inits = {"freestream": None,
                     "own_field_raster": ("raster", resample(full, R, R).astype(np.float32)),
                     "own_field_cells": ("cells",
                                         np.ascontiguousarray(full.transpose(1, 2, 0)).ravel()),
                     "own_field_cells_anisotropic": ("moments", moments.ravel())}
            if pred is not None:
                inits["network"] = ("raster", pred[j])
            res[i] = {}
            for name, init in inits.items():
                with Simulation([p.replace(first_sampling_step=0, n_step=a.max_steps)]) as sim:
                    if init is None:
                        sim.initial_conditions()
                    elif init[0] == "raster":
                        sim.init_from_fields(init[1][None])
                    elif init[0] == "cells":
                        sim.init_from_cell_fields(init[1])
                    else:
                        sim.init_from_cell_moments(init[1])
                    prev = (sim.binned_cells(B, B, 0), sim.state(0).sample_step)
                    hist = []
                    for _ in range(a.max_steps // a.check_every):
                        sim.run(a.check_every)
                        cur = (sim.binned_cells(B, B, 0), sim.state(0).sample_step)
                        e = field_error(moments_to_fields(cur[0] - prev[0], cur[1] - prev[1], k,
                                                          nc), ref)
                        hist.append([round(x, 4) for x in e])
                        prev = cur
                res[i][name] = hist
                _log(f"case {i} {name}: worst channel per window "
                     + " ".join(f"{max(e):.3f}" for e in hist))
        out["seeding"] = res
    elif a.phase == "batchref":
        # the case-by-case references of `--phase cases` (seed 9999, the average over the last
        # `--ref-sample` of `--ref-steps` steps) for the cases in `--cases`, run as one batch:
        # each case draws the same random numbers alone or in a batch, and its particles do not
        # depend on when sampling starts
        sel = parse_cases(a.cases, len(test))
        F = a.ref_steps - a.ref_sample
        cs = [test[i].replace(seed=9999, first_sampling_step=F - 1, n_step=a.ref_steps)
              for i in sel]
        t0 = time.perf_counter()
        with Simulation(cs) as sim:
            sim.initial_conditions()
            for done in range(0, a.ref_steps, 1000):
                sim.run(min(1000, a.ref_steps - done))
                _log(f"  references: step {done + 1000} of {a.ref_steps}")
            for j, i in enumerate(sel):
                assert sim.state(j).sample_step == a.ref_sample
                ref = moments_to_fields(sim.binned_cells(B, B, j), a.ref_sample, sim.constants(j),
                                        cells_per_bin(test[i], B, B))
                np.save(os.path.join(a.state, f"ref_{i}.npy"), ref)
        out["ref_s"] = time.perf_counter() - t0
        out["cases_done"] = sel
    elif a.phase == "batched":
        # all cases in one pool from the network's fields (or the freestream fill), each retired
        # once it is steady against its case-by-case reference
        refs = [np.load(os.path.join(a.state, f"ref_{i}.npy")) for i in range(len(test))]
        init = None
        if a.kind == "warm":
            init, out["predict_s"] = predict()
        done_at, t_init, gpu, iv = retiring_batch(test, refs, a.tol, a.rule, a.check_every,
                                                  a.max_steps, B, init_fields=init, log=_log)
        out.update({"kind": a.kind, "rule": a.rule, "tol": a.tol, "max_steps": a.max_steps,
                    "steps": done_at, "init_s": t_init, "steps_s": gpu, "intervals": iv,
                    "wall_s": out.get("predict_s", 0.0) + t_init + gpu})
        _save_json(os.path.join(a.state, f"batched_{a.kind}_{a.rule}_{a.tol}.json"), out)
        _log(f"batched {a.kind} ({a.rule}, {a.tol}): {sum(x is not None for x in done_at)} of "
             f"{len(test)} steady, {out['wall_s']:.1f} s")
    elif a.phase == "ceiling":
        # the warm start's ceiling: every case seeded from its own converged field, cell by cell
        # with a temperature per component, the best any prediction of local Maxwellians could
        # do. One batched reference run (seed 9999, `--ref-steps`, sampled over the last
        # `--ref-sample`) gives both the block references (saved as ref_<i>.npy for the
        # `batched` phase) and the seeding moments; then the batch runs from those moments,
        # each case retired once steady by `--rule` at `--tol`.
        F = a.ref_steps - a.ref_sample
        cs = [p.replace(seed=9999, first_sampling_step=F - 1, n_step=a.ref_steps) for p in test]
        t0 = time.perf_counter()
        refs, moments = [], []
        with Simulation(cs) as sim:
            sim.initial_conditions()
            for done in range(0, a.ref_steps, 1000):
                sim.run(min(1000, a.ref_steps - done))
                _log(f"  references: step {done + 1000} of {a.ref_steps}")
            for i, p in enumerate(test):
                assert sim.state(i).sample_step == a.ref_sample
                k = sim.constants(i)
                refs.append(moments_to_fields(sim.binned_cells(B, B, i), a.ref_sample, k,
                                              cells_per_bin(p, B, B)))
                np.save(os.path.join(a.state, f"ref_{i}.npy"), refs[-1])
                moments.append(cell_moments(sim.cells(i), a.ref_sample, k))
        out["ref_s"] = time.perf_counter() - t0
        _log(f"references: {out['ref_s']:.1f} s")
        done_at, t_init, gpu, iv = retiring_batch(test, refs, a.tol, a.rule, a.check_every,
                                                  a.max_steps, B,
                                                  init_moments=np.concatenate(moments), log=_log)
        out.update({"kind": "ceiling", "rule": a.rule, "tol": a.tol, "max_steps": a.max_steps,
                    "steps": done_at, "init_s": t_init, "steps_s": gpu, "intervals": iv,
                    "wall_s": t_init + gpu})
        _save_json(os.path.join(a.state, f"batched_ceiling_{a.rule}_{a.tol}.json"), out)
        _log(f"ceiling ({a.rule}, {a.tol}): {sum(x is not None for x in done_at)} of "
             f"{len(test)} steady, {out['wall_s']:.1f} s")
    elif a.phase == "table":
        # the batched runs of this `--state` (plain, warm, ceiling: `batched`/`ceiling` phases)
        # against the plain runs case by case: each case alone for the steps the plain batch
        # needed for it (a case draws the same numbers alone or in a batch), at its step time
        # alone (`retime`), plus its own fill
        rt = json.load(open(os.path.join(a.state, "retime.json")))
        plain = json.load(open(os.path.join(a.state, f"batched_plain_{a.rule}_{a.tol}.json")))
        cap = plain["max_steps"]
        need = [n or cap for n in plain["steps"]]
        case_by_case = sum(n * t for n, t in zip(need, rt["step_s"])) + sum(rt["init_s"])
        rows = {"plain, case by case": {"wall_s": case_by_case, "speedup": 1.0,
                                        "median_steps": float(np.median(need))}}
        for kind in ("plain", "warm", "ceiling"):
            path = os.path.join(a.state, f"batched_{kind}_{a.rule}_{a.tol}.json")
            if not os.path.exists(path):
                continue
            r = json.load(open(path))
            steps = [n or cap for n in r["steps"]]
            rows[f"{kind}, one batch, retiring"] = {
                "wall_s": r["wall_s"], "speedup": case_by_case / r["wall_s"],
                "median_steps": float(np.median(steps)),
                "steady": sum(n is not None for n in r["steps"])}
            if kind != "plain":
                per = [(n0 * t) / (n1 * t) for n0, n1, t in zip(need, steps, rt["step_s"])]
                rows[f"{kind}, one batch, retiring"]["per_case_step_ratio_median"] = \
                    float(np.median(per))
        out.update({"rule": a.rule, "tol": a.tol, "rows": rows})
        _save_json(os.path.join(a.state, f"table_{a.rule}_{a.tol}.json"), out)
    elif a.phase == "replay":
        # the plain runs case by case, timed with this build: each case by itself from the
        # freestream fill, sampling every step, for the steps its case-by-case run (`--phase
        # cases`) needed to be steady by `--rule` at `--tol` (`--max-steps` where it never was);
        # the runs are deterministic, so the same steps give the same fields
        rows = []
        for i, p in enumerate(test):
            c = json.load(open(os.path.join(a.state, f"case_{i}.json")))
            n = c["plain"]["steady"][f"tol_{a.tol}"][a.rule] or a.max_steps
            with Simulation([p.replace(first_sampling_step=0, n_step=n)]) as sim:
                t0 = time.perf_counter()
                sim.initial_conditions()
                t_init = time.perf_counter() - t0
                t0 = time.perf_counter()
                for done in range(0, n, 1000):
                    sim.run(min(1000, n - done))
                t_steps = time.perf_counter() - t0
            rows.append({"case": i, "steps": n, "init_s": t_init, "steps_s": t_steps})
            _log(f"  case {i}: {n} steps in {t_steps:.2f} s")
        out.update({"rule": a.rule, "tol": a.tol, "replay": rows,
                    "wall_s": sum(r["init_s"] + r["steps_s"] for r in rows)})
        _save_json(os.path.join(a.state, f"replay_plain_{a.rule}_{a.tol}.json"), out)
        _log(f"replay: plain case by case {out['wall_s']:.1f} s")
    elif a.phase == "retime":
        # every case's step time run by itself with this build, on a flow developed from the
        # network's fields (from the freestream fill without weights) and sampled every step
        # (the regime of the case-by-case runs), and its freestream fill
        pred = predict()[0] if a.weights and os.path.exists(a.weights) else None
        step_s, init_s = [], []
        for i, p in enumerate(test):
            with Simulation([p.replace(first_sampling_step=0, n_step=10 ** 6)]) as sim:
                t0 = time.perf_counter()
                sim.initial_conditions()
                init_s.append(time.perf_counter() - t0)
                if pred is not None:
                    sim.init_from_fields(pred[i][None])
                sim.run(a.check_every)
                t0 = time.perf_counter()
                sim.run(a.retime_steps)
                step_s.append((time.perf_counter() - t0) / a.retime_steps)
            _log(f"  case {i}: {1e3 * step_s[-1]:.3f} ms per step")
        out["step_s"] = step_s
        out["init_s"] = init_s
        _save_json(os.path.join(a.state, "retime.json"), out)
        _log(f"retime: {1e3 * sum(step_s):.2f} ms per step summed over the cases")
    elif a.phase == "summary":
        cases = [json.load(open(os.path.join(a.state, f"case_{i}.json")))
                 for i in range(len(test)) if os.path.exists(os.path.join(a.state, f"case_{i}.json"))]
        out["cases_found"] = len(cases)
        cap = a.max_steps
        summ = {"max_steps": cap, "ref_steps": cases[0]["ref_steps"] if cases else None,
                "ref_halves_error_median": float(np.median(
                    [max(c["ref_halves_error"]) for c in cases])) if cases else None}
        for tol in (0.02, 0.03, 0.05):
            for rule in ("hindsight", "second_half"):
                t = {}
                for kind in ("plain", "warm"):
                    ns = [c[kind]["steady"][f"tol_{tol}"][rule] for c in cases]
                    t[kind + "_steady"] = sum(x is not None for x in ns)
                    t[kind + "_median_steps"] = float(np.median([x or cap for x in ns]))
                    t[kind + "_case_by_case_s"] = float(sum(
                        c[kind]["init_s"] + c[kind]["step_s"] * (x or cap)
                        for c, x in zip(cases, ns)))
                t["speedup"] = t["plain_case_by_case_s"] / t["warm_case_by_case_s"]
                t["capped_cases"] = sum(
                    c["plain"]["steady"][f"tol_{tol}"][rule] is None
                    or c["warm"]["steady"][f"tol_{tol}"][rule] is None for c in cases)
                summ[f"tol_{tol}_{rule}"] = t
        # the retiring batches (`--phase batched`) against the plain runs case by case; with
        # `--phase retime`, the plain runs' step times are also scaled to this build (each case
        # by its warm run's step time now over then)
        retime = os.path.join(a.state, "retime.json")
        now = json.load(open(retime))["step_s"] if os.path.exists(retime) else None
        by_case = {c["case"]: c for c in cases}
        for name in sorted(os.listdir(a.state)):
            if not (name.startswith("batched_") and name.endswith(".json")) or \
                    len(cases) != len(test):
                continue
            b = json.load(open(os.path.join(a.state, name)))
            key = f"tol_{b['tol']}"
            ns = [by_case[i]["plain"]["steady"][key][b["rule"]] for i in range(len(test))]
            plain = summ[f"{key}_{b['rule']}"]["plain_case_by_case_s"]
            row = {"steady": sum(x is not None for x in b["steps"]),
                   "median_steps": float(np.median([x or b["max_steps"] for x in b["steps"]])),
                   "wall_s": b["wall_s"], "plain_case_by_case_s": plain,
                   "speedup": plain / b["wall_s"]}
            rp = os.path.join(a.state, f"replay_plain_{b['rule']}_{b['tol']}.json")
            if os.path.exists(rp):
                replayed = json.load(open(rp))["wall_s"]
                row.update({"plain_case_by_case_s_replayed": replayed,
                            "speedup_replayed": replayed / b["wall_s"]})
            if now is not None:
                plain_now = float(sum(
                    by_case[i]["plain"]["init_s"] + (x or cap) * by_case[i]["plain"]["step_s"]
                    * now[i] / by_case[i]["warm"]["step_s"] for i, x in enumerate(ns)))
                row.update({"plain_case_by_case_s_this_build": plain_now,
                            "speedup_this_build": plain_now / b["wall_s"]})
            summ[name[:-5]] = row
        out["summary"] = summ
    print(json.dumps(out), flush=True)


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
    e = sub.add_parser("evaluate", help="held-out sweep: field error and time to steady state")
    e.add_argument("--phase",
                   choices=["ref", "cases", "summary", "seeding", "batchref", "batched",
                            "replay", "retime", "ceiling", "table"],
                   required=True)
    e.add_argument("--cases", default=None,
                   help="case indices for --phase cases (e.g. 0-15 or 3,7); default all")
    e.add_argument("--warm-only", action="store_true",
                   help="--phase cases: rerun only the warm-started runs (new weights) against "
                        "the references and freestream runs an earlier run saved in --state")
    e.add_argument("--state", default="sweep_state")
    e.add_argument("--seed", type=int, default=0)
    e.add_argument("--raster", type=int, default=256)
    e.add_argument("--weights", default="warmstart_weights.npy")
    e.add_argument("--test-cases", type=int, default=64)
    e.add_argument("--ref-steps", type=int, default=20000)
    e.add_argument("--ref-sample", type=int, default=10000)
    e.add_argument("--bins", type=int, default=50)
    e.add_argument("--tol", type=float, default=0.02)
    e.add_argument("--check-every", type=int, default=500)
    e.add_argument("--max-steps", type=int, default=20000)
    e.add_argument("--kind", choices=["warm", "plain"], default="warm",
                   help="--phase batched: start from the network's fields or the freestream fill")
    e.add_argument("--rule", choices=["hindsight", "second_half"], default="hindsight",
                   help="--phase batched: the steady test that retires a case")
    e.add_argument("--retime-steps", type=int, default=1000)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        train_experiment(a)
    else:
        evaluate_experiment(a)


if __name__ == "__main__":
    main()
