"""Field statistics shared by the golden-field script and the field-parity tests.

Fields follow Printing's conversion (src/Sampling.cpp:152-196): n = W CountNP / SampleStep / DV,
U = CountV / CountNP, T_m = m (CountVV_mm / CountNP - U_m^2) / k, T = mean of the three.
Standard errors come from batch means: the sampling window is cut into B batches, each batch's
field is computed from the difference of the cumulative sums, and the error of the window mean
is std(batches) / sqrt(B) (consecutive samples are strongly correlated, so the naive
per-sample variance would understate it).
"""
import math
import warnings

import numpy as np

KB = 1.380662e-23
NA = 6.022045e23


def constants_for(p):
    m = p.molar_mass / NA
    dx = (p.x2 - p.x1) / p.nx
    dy = (p.y2 - p.y1) / p.ny
    dv = dx * dy * p.dz
    return {"m": m, "dv": dv, "w": dv * p.n_free / p.npc_free}


def fields(cells, nsamples, k):
    """cells: [M, 17] sums over nsamples steps -> [M, 4] (n, ux, uy, T); NaN where empty."""
    np_ = cells[:, 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        n = k["w"] * np_ / nsamples / k["dv"]
        u = cells[:, 1:4] / np_[:, None]
        t = k["m"] * (cells[:, [4, 8, 12]] / np_[:, None] - u * u) / KB
    out = np.stack([n, u[:, 0], u[:, 1], t.mean(1)], axis=1)
    out[np_ == 0, 1:] = np.nan
    return out


def batch_fields(snaps, per_batch, k):
    """snaps: [B+1, M, 17] cumulative sums at batch boundaries -> (mean [M,4], err [M,4])."""
    return multi_batch_fields([snaps], per_batch, k)


def multi_batch_fields(runs, per_batch, k):
    """Several independent runs of the same case over the same window, each given as [B+1, M, 17]
    cumulative sums at its batch boundaries: the field over all their samples and its standard
    error from the batch means of every run (batches of different runs are independent)."""
    batches = np.array([fields(s[i + 1] - s[i], per_batch, k)
                        for s in runs for i in range(len(s) - 1)])
    total = sum(s[-1] - s[0] for s in runs)
    mean = fields(total, per_batch * len(batches), k)
    with np.errstate(invalid="ignore"), warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)   # cells empty in every batch
        err = np.nanstd(batches, axis=0, ddof=1) / math.sqrt(len(batches))
    return mean, err


def compare(a, ea, b, eb, gas, dense=None):
    """Per field: the relative L2 difference, the part of it that the two runs' sampling noise
    explains (E|d|^2 = |bias|^2 + sum of variances), the remaining systematic part, and the
    fractions of cells beyond 3 and 5 combined standard errors. n is compared over the `gas`
    cells; velocity and temperature, which a handful of molecules cannot define, over the
    `dense` ones (default: the same)."""
    res = {}
    names = ["n", "ux", "uy", "T"]
    dense = gas if dense is None else dense
    for i, name in enumerate(names):
        ok = (gas if i == 0 else dense) & np.isfinite(a[:, i]) & np.isfinite(b[:, i]) \
            & np.isfinite(ea[:, i]) & np.isfinite(eb[:, i])
        d = a[ok, i] - b[ok, i]
        s = np.sqrt(ea[ok, i] ** 2 + eb[ok, i] ** 2)
        z = d / np.where(s > 0, s, np.inf)
        if name in ("ux", "uy"):  # relative to the speed, as uy alone is near zero
            ref = np.hypot(b[ok, 1], b[ok, 2])
        else:
            ref = b[ok, i]
        norm = np.linalg.norm(ref)
        l2 = float(np.linalg.norm(d) / norm)
        noise = float(np.linalg.norm(s) / norm)
        res[name] = {"l2": l2, "noise_l2": noise,
                     "bias_l2": float(math.sqrt(max(0.0, l2 * l2 - noise * noise))),
                     "beyond3": float(np.mean(np.abs(z) > 3)),
                     "beyond5": float(np.mean(np.abs(z) > 5)),
                     "cells": int(ok.sum())}
    return res
