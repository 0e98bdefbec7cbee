"""Sweep benchmark: the wall-clock of a 64-case sweep to steady fields within 2% L2, warm-started
by the learned network against plain freestream runs, on one MI355X.

The held-out cases are drawn from the same ranges as the training cases (seed + 7). The phases
keep their results in `--state`, so each fits one bounded GPU session; each prints one JSON
line. The protocol, the phases and their results are described in profiles/warmstart_sweep.md;
the building blocks (case draws, field errors, steady tests, retiring batches, reference runs)
are in dsmc_erlend_amd.sweep.

    python -m dsmc_erlend_amd.sweep train --weights w.npy
    python sweep_bench.py --phase ref --state DIR --weights w.npy
    python sweep_bench.py --phase batchref --cases 0,2,...,62 --state DIR
    python sweep_bench.py --phase batched --kind warm --state DIR --weights w.npy
    python sweep_bench.py --phase table --state DIR
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from dsmc_erlend_amd import Simulation  # noqa: E402
from dsmc_erlend_amd.surrogate import Surrogate  # noqa: E402
from dsmc_erlend_amd.sweep import (_log, case_fields, cell_moments, cells_per_bin,  # noqa: E402
                                   draw_cases, field_error, freestream_fields, moments_to_fields,
                                   reference_fields, resample, retiring_batch, snapshot_run,
                                   steady_steps, window_field)


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


def evaluate(a):
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
    e = argparse.ArgumentParser(prog="python sweep_bench.py")
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
    evaluate(e.parse_args(argv))


if __name__ == "__main__":
    main()
