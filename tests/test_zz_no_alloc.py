"""dsmc_get_binned_cells_noalloc, the entry point that allocates nothing, against the
allocating dsmc_get_binned_cells through the C-ABI: bit-identical block sums on two calls in a
row (the second smaller, so stale contents of the shared staging array would show), on two
contexts sharing one GPU, and the size limit."""
import ctypes as C

import numpy as np
import pytest

from dsmc_erlend_amd import DsmcError, Simulation, default_params, load_library
from dsmc_erlend_amd._abi import E_ARGUMENT

pytestmark = pytest.mark.gpu


def noalloc(sim, case, bx, by):
    """The entry point called directly through ctypes; the output starts as NaN."""
    out = np.full((bx, by, 17), np.nan)
    lib = load_library()
    rc = lib.dsmc_get_binned_cells_noalloc(sim.h, case, bx, by,
                                           out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_noalloc_binned_cells_equal_the_allocating_entry():
    cases = [default_params(nx=150, ny=130, first_sampling_step=0, case_id=0),
             default_params(nx=60, ny=45, wall_model=0, attack_angle_deg=12.0,
                            first_sampling_step=0, case_id=1)]
    a = Simulation(cases)
    a.initial_conditions()
    a.run(25)
    b = Simulation(default_params(nx=40, ny=40, first_sampling_step=0, case_id=5))
    b.initial_conditions()
    b.run(10)
    # two calls in a row, the second smaller; then the other context, then the first again
    for sim, case, bx, by in ((a, 0, 128, 128), (a, 1, 9, 7), (b, 0, 40, 40), (a, 0, 3, 2),
                              (a, 1, 60, 45)):
        rc, got = noalloc(sim, case, bx, by)
        assert rc == 0, sim.lib.dsmc_last_error(sim.h)
        want = sim.binned_cells(bx, by, case)
        np.testing.assert_array_equal(got, want)
        assert got[..., 0].sum() == sim.cells(case)[:, 0].sum()
    # the sums keep moving with the run: the staging holds nothing over from the last call
    a.run(3)
    rc, got = noalloc(a, 1, 9, 7)
    assert rc == 0
    np.testing.assert_array_equal(got, a.binned_cells(9, 7, 1))
    # the Python view of the same entry point
    np.testing.assert_array_equal(a.binned_cells_noalloc(20, 10, 0), a.binned_cells(20, 10, 0))


def test_noalloc_binned_cells_rejects_sizes_beyond_its_staging():
    s = Simulation(default_params(nx=150, ny=130, first_sampling_step=0))
    s.initial_conditions()
    s.run(2)
    rc, _ = noalloc(s, 0, 129, 128)          # 16512 blocks > 16384
    assert rc == E_ARGUMENT and b"16384" in s.lib.dsmc_last_error(s.h)
    rc, _ = noalloc(s, 0, 151, 10)           # more blocks than cells along x
    assert rc == E_ARGUMENT
    rc, _ = noalloc(s, 1, 10, 10)            # no such case
    assert rc == E_ARGUMENT
    with pytest.raises(DsmcError):
        s.binned_cells_noalloc(129, 128)
    # the context still works after the rejected calls
    np.testing.assert_array_equal(s.binned_cells_noalloc(128, 128), s.binned_cells(128, 128))
