"""The GPU's wall models, inflow and collisions on fed random draws, against the reference's own
functions (oracle/_ref, compiled from its sources) on the same draws.

Both sides consume one sequential stream: the reference's MWC generator (rng/BRNG.cxx:17-34)
from a fixed seed, which the oracle reproduces bit for bit (tests/test_oracle_pins.py), is fed
to the GPU through dsmc_set_draws while the reference draws it itself. The GPU starts from the
reference's particles with FP32 velocities, so the only differences left are the GPU's FP32
velocity stores and an ulp of the FP64 log/cos/sin/pow/hypot of the GPU's math library against
glibc: positions agree to 1e-12 m, velocities to 1e-6 relative, counts exactly. Both sides also
consume exactly the same number of draws: the reference's next draw is the first one the GPU
left unread (a wall model, a count or a collision test that took another path would shift it).
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as L
from dsmc_erlend_amd import DsmcError, Simulation, default_params
from dsmc_erlend_amd._abi import E_CAPACITY, WALL_CLL, WALL_DIFFUSE, WALL_SPECULAR

pytestmark = pytest.mark.gpu
SEED = (521288629, 362436069)
VTOL = 1e-6


@pytest.fixture(autouse=True)
def _ref():
    if not L.ref_available():
        pytest.skip("the reference was not compiled here (make -C oracle ref)")


def mwc(n, seed=SEED):
    d = L.Draws("mwc", *seed)
    return np.array([d.next() for _ in range(n)])


def f32(a):
    return np.asarray(a, np.float32)


def start(p, parts, draws):
    """The GPU context and the reference on the same particles (FP32 velocities) and draws."""
    x, y, vx, vy, vz = parts
    vx, vy, vz = f32(vx), f32(vy), f32(vz)
    g = Simulation(p)
    g.set_particles(x, y, vx, vy, vz)
    g.set_draws(draws)
    r = L.ref()
    assert r.ref_configure(p) == 0
    r.ref_clear_cells()
    r.ref_set_seed(*SEED)
    assert L.ref_set_particles(x, y, vx.astype(np.float64), vy.astype(np.float64),
                               vz.astype(np.float64)) == 0
    return g, r


def assert_same_consumption(g, r, draws):
    used = g.draws_used()
    assert 0 < used < draws.size
    assert r.ref_brng() == draws[used]


def assert_close(gp, rp, order=None):
    gx, gy, gvx, gvy, gvz = gp
    rx, ry, rvx, rvy, rvz = rp if order is None else [a[order] for a in rp]
    assert gx.size == rx.size
    assert np.max(np.abs(gx - rx), initial=0) < 1e-12
    assert np.max(np.abs(gy - ry), initial=0) < 1e-12
    for a, b in ((gvx, rvx), (gvy, rvy), (gvz, rvz)):
        assert np.allclose(a.astype(np.float64), b, rtol=VTOL, atol=1e-3)


def gas_near_box(n, rng, half=5.0, box=3.0, speed=800.0):
    x = rng.uniform(-half, half, 4 * n); y = rng.uniform(-half, half, 4 * n)
    keep = ~((np.abs(x) < box) & (np.abs(y) < box))
    x, y = x[keep][:n], y[keep][:n]
    v = rng.normal(0, speed, (3, x.size))
    return x, y, v[0], v[1], v[2]


@pytest.mark.parametrize("wall,an,at", [(WALL_DIFFUSE, 0.0, 0.0), (WALL_CLL, 0.8, 0.8),
                                        (WALL_CLL, 0.3, 0.7), (WALL_SPECULAR, 0.0, 0.0)])
def test_move_on_fed_draws_matches_reference(wall, an, at):
    """MoveParticles (src/Boundary.cpp:284-328) with DiffuseScattering and CLLScattering
    (:205-282): impacts, the particles after them, drag and the draws consumed."""
    p = default_params(box_width=6.0, box_height=6.0, wall_model=wall, acc_n=an, acc_t=at,
                       dt=5e-4)
    draws = mwc(200000)
    g, r = start(p, gas_near_box(20000, np.random.default_rng(20 + wall)), draws)
    g.move_particles()
    assert r.ref_move() == 0
    assert_close(g.particles(), L.ref_particles())
    imp, hits = ctypes.c_double(), ctypes.c_longlong()
    r.ref_drag(ctypes.byref(imp), ctypes.byref(hits))
    s = g.state()
    assert s.wall_hits == hits.value > 1000
    assert abs(s.drag_impulse_total - imp.value) <= 1e-9 * abs(imp.value)
    if wall == WALL_SPECULAR:
        assert g.draws_used() == 0 and r.ref_brng() == draws[0]
    else:
        assert_same_consumption(g, r, draws)


def test_boundary_on_fed_draws_matches_reference():
    """BoundaryConditions (src/Boundary.cpp:94-203): the four strips' counts and new particles
    (GenerateNewParticles, :60-92) and the removal. The reference removes by swap-with-last and
    the GPU keeps order, so the pools are compared as sets (sorted by position)."""
    p = default_params(nx=50, ny=50, npc_free=10, attack_angle_deg=20.0, capacity_factor=8.0)
    rng = np.random.default_rng(5)
    n = 30000
    parts = (rng.uniform(-20.5, 20.5, n), rng.uniform(-20.5, 20.5, n), *rng.normal(0, 400, (3, n)))
    draws = mwc(400000)
    g, r = start(p, parts, draws)
    g.boundary_conditions()
    r.ref_boundary()
    gp, rp = g.particles(), L.ref_particles()
    assert gp[0].size == rp[0].size
    assert_close([a[np.lexsort((gp[1], gp[0]))] for a in gp], rp, np.lexsort((rp[1], rp[0])))
    assert g.state().last_injected > 0
    assert_same_consumption(g, r, draws)


def dense_gas(p, per_cell, rng, t=200.0):
    m = 0.040 / 6.022045e23
    s = np.sqrt(1.380662e-23 * t / m)
    dx = (p.x2 - p.x1) / p.nx; dy = (p.y2 - p.y1) / p.ny
    k = np.repeat(np.arange(p.nx), p.ny * per_cell)
    l = np.tile(np.repeat(np.arange(p.ny), per_cell), p.nx)
    x = p.x1 + (k + rng.uniform(0.05, 0.95, k.size)) * dx
    y = p.y1 + (l + rng.uniform(0.05, 0.95, k.size)) * dy
    v = rng.normal(0, s, (3, x.size))
    return x, y, v[0], v[1], v[2]


def test_collide_on_fed_draws_matches_reference():
    """CollideParticles (src/Collision.cpp:45-56): per cell the NTC count's draw, the candidate
    pairs, the acceptance tests and the isotropic scattering; collision counts and CountNC exact.
    The GPU's pool is sorted by cell; its particle i is the reference's perm[i]."""
    p = default_params(nx=10, ny=10, dt=2e-3, npc_free=40, first_sampling_step=0)
    draws = mwc(400000)
    g, r = start(p, dense_gas(p, 40, np.random.default_rng(4)), draws)
    g.set_draws([])                    # indexing draws nothing; the fed stream starts at collide
    g.indexing()
    _, perm = g.index()
    g.set_draws(draws)
    g.set_step(1)
    _, rperm = L.ref_index(p.nx * p.ny, perm.size)
    assert np.array_equal(rperm, perm)
    r.ref_set_step(1)
    nc = g.collide_particles()
    assert nc == r.ref_collide() > 100
    assert_close(g.particles(), L.ref_particles(), perm)
    assert np.array_equal(g.cells()[:, 16], L.ref_cells(p.nx * p.ny)[:, 16])
    assert_same_consumption(g, r, draws)


def test_fed_draws_that_run_out_fail():
    p = default_params(box_width=6.0, box_height=6.0, wall_model=WALL_DIFFUSE, dt=5e-4)
    g = Simulation(p)
    g.set_particles(*gas_near_box(5000, np.random.default_rng(1)))
    g.set_draws(mwc(10))
    with pytest.raises(DsmcError) as e:
        g.move_particles()
    assert e.value.code == E_CAPACITY
    with pytest.raises(DsmcError):
        g.set_draws([0.5, 1.0])        # draws lie in [0, 1)
