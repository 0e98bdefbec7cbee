"""Host logic of the sweep library and its benchmark (CPU only): case selection, blocks of cells, moments to fields,
and the time-to-steady criterion of the case-by-case protocol against closed forms."""
import types

import numpy as np
import pytest

from dsmc_erlend_amd.sweep import (KB, SteadyTest, bin_index, cell_moments, cells_per_bin,
                                   field_error, moments_to_fields, steady_steps, window_field)
from sweep_bench import parse_cases


def test_parse_cases():
    assert parse_cases("0-3,7,70", 64) == [0, 1, 2, 3, 7]
    assert parse_cases(None, 5) == [0, 1, 2, 3, 4]
    assert parse_cases("4", 3) == []


@pytest.mark.parametrize("n,nb", [(100, 50), (767, 50), (235, 50), (50, 50), (51, 7)])
def test_bins_cover_every_cell_once_and_stay_balanced(n, nb):
    b = bin_index(n, nb)
    assert b.min() == 0 and b.max() == nb - 1
    assert np.all(np.diff(b) >= 0)
    counts = np.bincount(b, minlength=nb)
    assert counts.sum() == n and counts.max() - counts.min() <= 1
    p = types.SimpleNamespace(nx=n, ny=n + 3)
    assert cells_per_bin(p, nb, nb).sum() == n * (n + 3)


K = types.SimpleNamespace(weight=1e12, dv=0.0016, molecule_mass=6.63e-26)


def sums_for(n_per_step, ux, uy, T, ncell, steps):
    """CELL sums [1, 1, 17] of `steps` steps with the given per-step density and a Maxwellian
    of velocity (ux, uy, 0) and temperature T (exact second moments)."""
    npart = n_per_step * K.dv * ncell / K.weight
    s = np.zeros((1, 1, 17))
    vt2 = KB * T / K.molecule_mass
    u = (ux, uy, 0.0)
    s[..., 0] = npart * steps
    for a in range(3):
        s[..., 1 + a] = npart * steps * u[a]
        s[..., 4 + 4 * a] = npart * steps * (u[a] ** 2 + vt2)
    return s


def test_moments_to_fields_recovers_the_maxwellian():
    s = sums_for(2.5e17, 900.0, -120.0, 350.0, ncell=12, steps=400)
    f = moments_to_fields(s, 400, K, 12)
    assert f[:, 0, 0] == pytest.approx([2.5e17, 900.0, -120.0, 350.0], rel=1e-12)


def snapshots(amp, tau, every, steps, ncell=12, n_inf=1e17, T=300.0):
    """Block sums every `every` steps of a run whose density relaxes as
    n(t) = n_inf (1 + amp r^t), r = exp(-1/tau), at fixed velocity and temperature."""
    snaps = [(0, np.zeros((1, 1, 17)), 0)]
    total = np.zeros((1, 1, 17))
    for t in range(1, steps + 1):
        total = total + sums_for(n_inf * (1 + amp * np.exp(-t / tau)), 0.0, 0.0, T, ncell, 1)
        if t % every == 0:
            snaps.append((t, total.copy(), t))
    ref = moments_to_fields(sums_for(n_inf, 0.0, 0.0, T, ncell, 1), 1, K, ncell)
    return snaps, ref


def window_deviation(amp, tau, F, N):
    """Relative density deviation of the average over steps F+1..N: the geometric series
    amp * sum r^t / (N - F), r = exp(-1/tau)."""
    r = np.exp(-1.0 / tau)
    return amp * r ** (F + 1) * (1 - r ** (N - F)) / (1 - r) / (N - F)


@pytest.mark.parametrize("amp,tau", [(0.5, 300.0), (-0.4, 800.0), (0.05, 50.0)])
def test_steady_steps_match_the_closed_form(amp, tau):
    every, steps = 100, 4000
    snaps, ref = snapshots(amp, tau, every, steps)
    tols = (0.02, 0.05)
    out, best, half = steady_steps(snaps, ref, K, 12, tols)
    grid = list(range(every, steps + 1, every))
    for tol in tols:
        # hindsight: the best F for an N is the last grid point before it (the transient only
        # decays), so N is steady once the last interval's average is within the tolerance
        want = next((N for N in grid if abs(window_deviation(amp, tau, N - every, N)) < tol),
                    None)
        want_half = next((N for N in grid
                          if abs(window_deviation(amp, tau, (N // every // 2) * every, N)) < tol),
                         None)
        assert out[f"tol_{tol}"]["hindsight"] == want
        assert out[f"tol_{tol}"]["second_half"] == want_half
    for (N, e), (N2, e2) in zip(best, half):
        assert N == N2 and e <= e2 + 1e-15


@pytest.mark.parametrize("rule", ["hindsight", "second_half"])
def test_online_steady_test_agrees_with_the_offline_one(rule):
    """The check that retires a case in a batch (SteadyTest, fed each snapshot as it arrives and
    keeping only the snapshots its rule can use) fires first at the step the offline search over
    the whole run finds, with the same error at every check."""
    snaps, ref = snapshots(0.5, 300.0, 100, 3000)
    for tol in (0.02, 0.05):
        out, best, half = steady_steps(snaps, ref, K, 12, (tol,))
        errs = best if rule == "hindsight" else half
        t = SteadyTest(ref, K, 12, rule, max_starts=len(snaps))
        assert t.add(snaps[0][1], snaps[0][2]) is None
        first = None
        for b in range(1, len(snaps)):
            e = t.add(snaps[b][1], snaps[b][2])
            assert e == pytest.approx(errs[b - 1][1], rel=1e-12)
            if first is None and e < tol:
                first = snaps[b][0]
        assert first is not None and first == out[f"tol_{tol}"][rule]
        assert len(t.snaps) <= (len(snaps) + 1) // 2 + 1 or rule == "hindsight"


def test_hindsight_starts_are_capped():
    """With max_starts m, the hindsight test tries the windows starting at the last m snapshots
    only: its error is the offline minimum over those starts, and it keeps m + 1 snapshots."""
    snaps, ref = snapshots(-0.4, 800.0, 100, 2500)
    m = 5
    t = SteadyTest(ref, K, 12, "hindsight", max_starts=m)
    t.add(snaps[0][1], snaps[0][2])
    for b in range(1, len(snaps)):
        e = t.add(snaps[b][1], snaps[b][2])
        want = min(max(field_error(window_field(snaps, a, b, K, 12), ref))
                   for a in range(max(0, b - m), b))
        assert e == pytest.approx(want, rel=1e-12)
        assert len(t.snaps) <= m + 1


def test_cell_moments_recover_an_anisotropic_maxwellian():
    """Per-cell (n, ux, uy, Tx, Ty, Tz) from exact CELL sums of a Maxwellian with a temperature
    per component; an empty cell gives zeros."""
    steps, npart = 250, 37.0
    u, T = (640.0, -75.0, 12.0), (180.0, 420.0, 900.0)
    cells = np.zeros((2, 17))
    cells[0, 0] = npart * steps
    for a in range(3):
        cells[0, 1 + a] = npart * steps * u[a]
        cells[0, 4 + 4 * a] = npart * steps * (u[a] ** 2 + KB * T[a] / K.molecule_mass)
    m = cell_moments(cells, steps, K)
    n = K.weight * npart / K.dv
    assert m[0] == pytest.approx([n, u[0], u[1], T[0], T[1], T[2]], rel=1e-9)
    assert np.all(m[1] == 0.0)


def test_field_error_of_a_uniform_offset():
    ref = np.ones((4, 3, 3))
    ref[1] = 2.0
    ref[2] = 0.0
    f = ref * 1.03
    e = field_error(f, ref)
    assert e == pytest.approx((0.03, 0.03, 0.03), rel=1e-9)


def test_sweep_bench_command_line():
    """sweep_bench.py's phases parse (the measurement driver lives beside bench.py)."""
    import sweep_bench
    with pytest.raises(SystemExit) as e:
        sweep_bench.main(["--help"])
    assert e.value.code == 0
    with pytest.raises(SystemExit) as e:
        sweep_bench.main(["--phase", "nonesuch"])
    assert e.value.code == 2
