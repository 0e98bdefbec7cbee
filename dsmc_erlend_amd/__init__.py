"""dsmc_erlend_amd - the MI355X (gfx950) DSMC per-step particle pipeline.

Python access to the C-ABI of include/dsmc_c.h (libdsmc_hip.so, built in-tree by
dsmc_erlend_amd.build). The class `Simulation` keeps the reference's step names
(src/Simulation.cpp:143-159): move_particles, boundary_conditions, indexing,
collide_particles, sampling, plus run() for the fused, graph-captured loop. There is no CPU
fallback: without the built library every call raises.
"""
import ctypes as C
import os

import numpy as np

from ._abi import (Constants, Params, State, Tuning, default_params, WALL_CLL, WALL_DIFFUSE,
                   WALL_SPECULAR, WALL_NAMES, STAGES, SCHEDULE_LARGE, SCHEDULE_SMALL)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsmc_hip.so")

__all__ = ["Simulation", "Params", "Constants", "State", "Tuning", "default_params", "DsmcError",
           "SCHEDULE_LARGE", "SCHEDULE_SMALL",
           "load_library", "WALL_CLL", "WALL_DIFFUSE", "WALL_SPECULAR", "WALL_NAMES"]


class DsmcError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"dsmc error {code}: {message}")
        self.code = code


_LIB = None
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)


def load_library(path=None):
    """Loads libdsmc_hip.so and declares every entry point of include/dsmc_c.h."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m dsmc_erlend_amd.build`")
    lib = C.CDLL(path)
    vp = C.c_void_p
    ctxpp = C.POINTER(C.c_void_p)
    sig = {
        "dsmc_params_default": (None, [C.POINTER(Params)]),
        "dsmc_create": (C.c_int, [C.POINTER(Params), C.c_int, ctxpp]),
        "dsmc_create_batch": (C.c_int, [C.POINTER(Params), C.c_int, C.c_int, ctxpp]),
        "dsmc_destroy": (None, [vp]),
        "dsmc_last_error": (C.c_char_p, [vp]),
        "dsmc_num_cases": (C.c_int, [vp]),
        "dsmc_get_params": (C.c_int, [vp, C.c_int, C.POINTER(Params)]),
        "dsmc_get_constants": (C.c_int, [vp, C.c_int, C.POINTER(Constants)]),
        "dsmc_get_state": (C.c_int, [vp, C.c_int, C.POINTER(State)]),
        "dsmc_init_freestream": (C.c_int, [vp]),
        "dsmc_init_from_fields": (C.c_int, [vp, _fp, C.c_int, C.c_int]),
        "dsmc_init_from_cell_fields": (C.c_int, [vp, _dp]),
        "dsmc_init_from_cell_moments": (C.c_int, [vp, _dp]),
        "dsmc_move": (C.c_int, [vp]),
        "dsmc_boundary": (C.c_int, [vp]),
        "dsmc_index": (C.c_int, [vp]),
        "dsmc_collide": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "dsmc_sample": (C.c_int, [vp]),
        "dsmc_step": (C.c_int, [vp, C.c_int]),
        "dsmc_retire_case": (C.c_int, [vp, C.c_int]),
        "dsmc_set_step": (C.c_int, [vp, C.c_int64]),
        "dsmc_set_draws": (C.c_int, [vp, _dp, C.c_size_t]),
        "dsmc_draws_used": (C.c_int, [vp, C.POINTER(C.c_size_t)]),
        "dsmc_profile_steps": (C.c_int, [vp, C.c_int, _dp]),
        "dsmc_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "dsmc_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "dsmc_autotune": (C.c_int, [vp, C.c_int, C.POINTER(Tuning)]),
        "dsmc_get_num_particles": (C.c_int, [vp, C.c_int, C.POINTER(C.c_size_t)]),
        "dsmc_get_particles": (C.c_int, [vp, C.c_int, _dp, _dp, _fp, _fp, _fp, C.c_size_t,
                                         C.POINTER(C.c_size_t)]),
        "dsmc_set_particles": (C.c_int, [vp, C.c_int, _dp, _dp, _fp, _fp, _fp, C.c_size_t]),
        "dsmc_get_index": (C.c_int, [vp, C.c_int, _u32p, C.c_size_t, _u32p, C.c_size_t]),
        "dsmc_get_cells": (C.c_int, [vp, C.c_int, _dp, C.c_size_t]),
        "dsmc_get_binned_cells": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, _dp]),
        "dsmc_get_binned_cells_noalloc": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, _dp]),
        "dsmc_get_fields": (C.c_int, [vp, C.c_int, _dp, C.c_size_t]),
        "dsmc_write_field": (C.c_int, [vp, C.c_int, C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path == LIB_PATH:
        _LIB = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(t)


class Simulation:
    """One context (one case, or a batch of independent cases sharing one particle pool)."""

    def __init__(self, params, device=0):
        self.lib = load_library()
        if isinstance(params, Params):
            params = [params]
        arr = (Params * len(params))(*params)
        self.params = list(params)
        self.ncases = len(params)
        h = C.c_void_p()
        rc = self.lib.dsmc_create_batch(arr, len(params), device, C.byref(h))
        self.h = h
        if rc:
            msg = self.lib.dsmc_last_error(h).decode() if h.value else "create failed"
            self.close()
            raise DsmcError(rc, msg)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.dsmc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc:
            raise DsmcError(rc, self.lib.dsmc_last_error(self.h).decode())

    def ncells(self, case=0):
        return self.params[case].nx * self.params[case].ny

    # --- setup ---
    def constants(self, case=0):
        k = Constants()
        self._check(self.lib.dsmc_get_constants(self.h, case, C.byref(k)))
        return k

    def state(self, case=0):
        s = State()
        self._check(self.lib.dsmc_get_state(self.h, case, C.byref(s)))
        return s

    def initial_conditions(self):
        self._check(self.lib.dsmc_init_freestream(self.h))

    def init_from_fields(self, fields):
        """fields: float32 array [ncases, 4, w, h] of (n, ux, uy, T)."""
        f = np.ascontiguousarray(fields, dtype=np.float32)
        if f.ndim == 3:
            f = f[None]
        self._check(self.lib.dsmc_init_from_fields(self.h, _p(f, _fp), f.shape[2], f.shape[3]))

    def _cell_values(self, values, nf):
        f = np.ascontiguousarray(values, dtype=np.float64).ravel()
        m = sum(self.ncells(i) for i in range(len(self.params)))
        if f.size != nf * m:
            raise ValueError(f"expected {nf} values for each of {m} cells, got {f.size}")
        return f

    def init_from_cell_fields(self, cell_fields):
        """Warm start from (n, ux, uy, T) per cell, cases concatenated, k*NY+l order."""
        f = self._cell_values(cell_fields, 4)
        self._check(self.lib.dsmc_init_from_cell_fields(self.h, _p(f, _dp)))

    def init_from_cell_moments(self, cell_moments):
        """Warm start from (n, ux, uy, Tx, Ty, Tz) per cell: a temperature per component."""
        f = self._cell_values(cell_moments, 6)
        self._check(self.lib.dsmc_init_from_cell_moments(self.h, _p(f, _dp)))

    # --- the step stages ---
    def move_particles(self):
        self._check(self.lib.dsmc_move(self.h))

    def boundary_conditions(self):
        self._check(self.lib.dsmc_boundary(self.h))

    def indexing(self):
        self._check(self.lib.dsmc_index(self.h))

    def collide_particles(self):
        nc = C.c_int64(0)
        self._check(self.lib.dsmc_collide(self.h, C.byref(nc)))
        return nc.value

    def sampling(self):
        self._check(self.lib.dsmc_sample(self.h))

    def run(self, nsteps):
        self._check(self.lib.dsmc_step(self.h, int(nsteps)))

    def retire(self, case):
        """Stops case `case` of the batch from the next step on; its sums stay readable."""
        self._check(self.lib.dsmc_retire_case(self.h, int(case)))

    def set_step(self, step):
        self._check(self.lib.dsmc_set_step(self.h, int(step)))

    def set_draws(self, u):
        """Fed draws (dsmc_set_draws): the single-stage move, boundary and collide calls take
        their uniforms from `u` in the reference's order; an empty `u` turns them off."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._check(self.lib.dsmc_set_draws(self.h, _p(u, _dp) if u.size else None, u.size))

    def draws_used(self):
        n = C.c_size_t(0)
        self._check(self.lib.dsmc_draws_used(self.h, C.byref(n)))
        return n.value

    def profile(self, nsteps):
        """Mean milliseconds per step of each stage (HIP events, eager launches)."""
        out = np.zeros(len(STAGES))
        self._check(self.lib.dsmc_profile_steps(self.h, int(nsteps), _p(out, _dp)))
        return dict(zip(STAGES, out.tolist()))

    def tuning(self):
        """(schedule, chunk, particles_per_thread) of the fused step."""
        t = Tuning()
        self._check(self.lib.dsmc_get_tuning(self.h, C.byref(t)))
        return t.schedule, t.chunk, t.particles_per_thread

    def set_tuning(self, schedule=None, chunk=None, particles_per_thread=None):
        s0, c0, p0 = self.tuning()
        t = Tuning(s0 if schedule is None else int(schedule), c0 if chunk is None else int(chunk),
                   p0 if particles_per_thread is None else int(particles_per_thread))
        self._check(self.lib.dsmc_set_tuning(self.h, C.byref(t)))

    def autotune(self, nsteps=20):
        """Times every admissible (schedule, chunk), then the particles per thread of the best
        large-schedule pair, over nsteps fused steps each (which advance the simulation); keeps
        and returns the fastest (schedule, chunk, particles_per_thread)."""
        t = Tuning()
        self._check(self.lib.dsmc_autotune(self.h, int(nsteps), C.byref(t)))
        return t.schedule, t.chunk, t.particles_per_thread

    # --- data ---
    def num_particles(self, case=0):
        n = C.c_size_t(0)
        self._check(self.lib.dsmc_get_num_particles(self.h, case, C.byref(n)))
        return n.value

    def particles(self, case=0):
        n = self.num_particles(case)
        x = np.zeros(n); y = np.zeros(n)
        vx = np.zeros(n, np.float32); vy = np.zeros(n, np.float32); vz = np.zeros(n, np.float32)
        got = C.c_size_t(0)
        self._check(self.lib.dsmc_get_particles(self.h, case, _p(x, _dp), _p(y, _dp),
                                                _p(vx, _fp), _p(vy, _fp), _p(vz, _fp), n,
                                                C.byref(got)))
        return x, y, vx, vy, vz

    def set_particles(self, x, y, vx, vy, vz, case=0):
        x = np.ascontiguousarray(x, np.float64); y = np.ascontiguousarray(y, np.float64)
        vx = np.ascontiguousarray(vx, np.float32); vy = np.ascontiguousarray(vy, np.float32)
        vz = np.ascontiguousarray(vz, np.float32)
        self._check(self.lib.dsmc_set_particles(self.h, case, _p(x, _dp), _p(y, _dp),
                                                _p(vx, _fp), _p(vy, _fp), _p(vz, _fp), x.size))

    def index(self, case=0, with_perm=True):
        m = self.ncells(case)
        counts = np.zeros(m, np.uint32)
        n = self.num_particles(case)
        perm = np.zeros(max(n, 1), np.uint32)
        self._check(self.lib.dsmc_get_index(self.h, case, _p(counts, _u32p), m,
                                            _p(perm, _u32p) if with_perm else None, n))
        return counts, perm[:n]

    def cells(self, case=0):
        m = self.ncells(case)
        out = np.zeros((m, 17))
        self._check(self.lib.dsmc_get_cells(self.h, case, _p(out, _dp), m))
        return out

    def binned_cells(self, bx, by, case=0):
        """CELL sums [bx, by, 17] added over blocks of cells (dsmc_get_binned_cells)."""
        out = np.zeros((bx, by, 17))
        self._check(self.lib.dsmc_get_binned_cells(self.h, case, int(bx), int(by), _p(out, _dp)))
        return out

    def binned_cells_noalloc(self, bx, by, case=0, out=None):
        """binned_cells through the entry point that allocates nothing on the device (its staging
        is a fixed per-GPU array of at most 16384 blocks); `out` [bx, by, 17] float64 may be
        given to reuse a host array."""
        if out is None:
            out = np.zeros((bx, by, 17))
        if out.shape != (bx, by, 17) or out.dtype != np.float64 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float64 array [bx, by, 17]")
        self._check(self.lib.dsmc_get_binned_cells_noalloc(self.h, case, int(bx), int(by),
                                                           _p(out, _dp)))
        return out

    def fields(self, case=0):
        m = self.ncells(case)
        out = np.zeros((m, 6))
        self._check(self.lib.dsmc_get_fields(self.h, case, _p(out, _dp), m))
        return out

    def write_field(self, directory, case=0):
        self._check(self.lib.dsmc_write_field(self.h, case, str(directory).encode()))
