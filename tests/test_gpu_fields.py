"""Sampled fields of the GPU pipeline against seeded runs of the reference's own code.

The golden fixtures (tests/golden/ref_fields_*.npz, written by tests/golden/make_ref_fields.py
with the reference compiled from its sources) hold per-cell mean fields (n, ux, uy, T) over a
sampling window, averaged over 16 independently seeded reference runs, and their batch-means
standard errors. The GPU runs the same case as 16 independently seeded cases in one batch (case
ids 0-15: their own Philox streams) over the same window, so the two fields differ by sampling
noise plus any systematic error. n is compared over every gas cell; ux, uy and T over the cells
holding at least 1% of the freestream density (in the near-vacuum wake right behind the box a cell
sees a few hundred molecules over the whole window, and its velocity and temperature are neither
defined nor Gaussian). Acceptance, per field (ux, uy relative to the speed): the relative L2
difference of the two fields <= 2% (noise included: at 16 x 8000 samples the two runs' combined
noise is 0.4-1.2% of the field), its systematic part (what the combined standard errors do not
explain) <= 1%, at most 0.6% of cells beyond 3 combined standard errors (Gaussian 0.27%) and at
most 0.05% beyond 5. A plausible mistake (a wrong wall temperature, a lost collision term, a
transposed velocity component) moves a whole region by many standard errors and fails all four.
The values reached are reported as a FieldParity warning, so they appear in the test summary.
"""
import json
import os
import warnings

import numpy as np
import pytest

from dsmc_erlend_amd import Simulation, default_params
from field_stats import compare, constants_for, multi_batch_fields

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GPU_RUNS = 16
MAX_L2, MAX_BIAS_L2, MAX_BEYOND3, MAX_BEYOND5 = 0.02, 0.01, 0.006, 0.0005


class FieldParity(UserWarning):
    """Field-parity figures reached (reported, not a failure)."""


def load(name):
    path = os.path.join(HERE, "golden", f"ref_fields_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated")
    z = np.load(path)
    meta = json.loads(str(z["meta"]))
    return z["mean"], z["err"], meta


@pytest.mark.parametrize("name", ["c1_specular", "c2_diffuse"])
def test_fields_match_seeded_reference_run(name):
    ref, ref_err, meta = load(name)
    pd = dict(meta["params"])
    p = default_params(**pd)
    cases = [p.replace(case_id=i) for i in range(GPU_RUNS)]
    sim = Simulation(cases)
    sim.initial_conditions()
    sim.run(meta["first_sampling"] + 1)
    per = meta["samples"] // meta["batches"]
    snaps = [[sim.cells(i)] for i in range(GPU_RUNS)]
    for _ in range(meta["batches"]):
        sim.run(per)
        for i in range(GPU_RUNS):
            snaps[i].append(sim.cells(i))
    mine, mine_err = multi_batch_fields([np.array(s) for s in snaps], per, constants_for(p))
    gas = np.isfinite(ref[:, 3]) & (ref[:, 0] > 0)
    dense = gas & (ref[:, 0] >= 0.01 * p.n_free)
    res = compare(mine, mine_err, ref, ref_err, gas, dense)
    print(name, json.dumps(res))
    warnings.warn(FieldParity(
        f"{name} ({meta.get('runs', 1)} reference runs, {GPU_RUNS} GPU runs, "
        f"{meta['samples']} samples each): " + "; ".join(
            f"{f} l2 {r['l2']:.4f} bias {r['bias_l2']:.4f} >3sd {r['beyond3']:.4f} "
            f">5sd {r['beyond5']:.4f}" for f, r in res.items())))
    for field, r in res.items():
        assert r["l2"] <= MAX_L2, (field, r)
        assert r["bias_l2"] <= MAX_BIAS_L2, (field, r)
        assert r["beyond3"] <= MAX_BEYOND3, (field, r)
        assert r["beyond5"] <= MAX_BEYOND5, (field, r)


def test_binned_cells_equal_host_binning_of_cells():
    """dsmc_get_binned_cells adds each case's CELL sums over blocks of cells on the device in the
    order a host pass over dsmc_get_cells would, so the two agree bit for bit; across a batch
    each case is binned on its own grid."""
    from dsmc_erlend_amd.sweep import bin_index
    cases = [default_params(nx=37, ny=23, first_sampling_step=0, case_id=0),
             default_params(nx=50, ny=64, first_sampling_step=0, attack_angle_deg=10.0,
                            wall_model=0, case_id=1)]
    with Simulation(cases) as sim:
        sim.initial_conditions()
        sim.run(40)
        for i, p in enumerate(cases):
            cells = sim.cells(i).reshape(p.nx, p.ny, 17)
            for bx, by in ((5, 7), (p.nx, p.ny), (1, 1), (16, 16)):
                got = sim.binned_cells(bx, by, i)
                kb, lb = bin_index(p.nx, bx), bin_index(p.ny, by)
                want = np.zeros((bx, by, 17))
                for k in range(p.nx):
                    for l in range(p.ny):
                        want[kb[k], lb[l]] += cells[k, l]
                np.testing.assert_array_equal(got, want)
        with pytest.raises(Exception):
            sim.binned_cells(0, 3, 0)
