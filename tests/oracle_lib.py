"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was built, to the
reference's own functions (oracle/_ref/libdsmc_ref.so). Test infrastructure only."""
import ctypes as C
import os

import numpy as np

from dsmc_erlend_amd._abi import Constants, Params, State

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libdsmc_ref.so")

MODE_REF, MODE_GPU = 0, 1
P_MOVE, P_INJ_COUNT, P_INJECT, P_COLLIDE, P_INIT_COUNT, P_INIT, P_WARM = 1, 2, 3, 4, 5, 6, 7

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_int64)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def _load_oracle():
    lib = C.CDLL(ORACLE_SO)
    vp = C.c_void_p
    sig = {
        "or_draws_new_mwc": (vp, [C.c_uint32, C.c_uint32]),
        "or_draws_new_philox": (vp, [C.c_uint32] * 5),
        "or_draws_new_array": (vp, [_dp, C.c_size_t]),
        "or_draws_free": (None, [vp]),
        "or_draws_next": (C.c_double, [vp]),
        "or_draws_used": (C.c_size_t, [vp]),
        "or_set_mwc_gaussian_order": (None, [C.c_int]),
        "or_philox4x32_10": (None, [_u32p, _u32p, _u32p]),
        "or_irand_uniform": (C.c_int, [C.c_int, vp]),
        "or_irand_poisson": (C.c_int, [C.c_double, vp]),
        "or_frand_rayleigh": (C.c_double, [C.c_double, vp]),
        "or_frand_gaussian": (C.c_double, [C.c_double, C.c_double, vp]),
        "or_v3rand_isotropic": (None, [_dp, vp]),
        "or_inside_box": (C.c_int, [C.c_double] * 6),
        "or_intersect_box": (C.c_int, [C.c_double] * 8 + [_dp, _dp, _dp]),
        "or_specular": (None, [_dp, C.c_double, C.c_double]),
        "or_diffuse": (None, [_dp] + [C.c_double] * 4 + [vp]),
        "or_cll": (C.c_int, [_dp] + [C.c_double] * 6 + [vp]),
        "or_params_default": (None, [C.POINTER(Params)]),
        "or_constants_of": (C.c_int, [C.POINTER(Params), C.POINTER(Constants)]),
        "or_create": (vp, [C.POINTER(Params), C.c_int, C.POINTER(C.c_int)]),
        "or_destroy": (None, [vp]),
        "or_set_mwc_seed": (None, [vp, C.c_uint32, C.c_uint32]),
        "or_init_freestream": (C.c_int, [vp]),
        "or_set_particles": (C.c_int, [vp, C.c_size_t, _dp, _dp, _dp, _dp, _dp]),
        "or_num_particles": (C.c_size_t, [vp]),
        "or_get_particles": (None, [vp, _dp, _dp, _dp, _dp, _dp]),
        "or_set_step": (None, [vp, C.c_int64]),
        "or_move": (C.c_int, [vp]),
        "or_boundary": (C.c_int, [vp]),
        "or_last_strip_counts": (None, [vp, _i64p]),
        "or_index": (C.c_int, [vp]),
        "or_get_index": (None, [vp, _u32p, _u32p]),
        "or_collide": (C.c_int, [vp, _i64p]),
        "or_sample": (C.c_int, [vp]),
        "or_step": (C.c_int, [vp, C.c_int]),
        "or_get_cells": (None, [vp, _dp]),
        "or_get_fields": (None, [vp, _dp]),
        "or_get_state": (None, [vp, C.POINTER(State)]),
        "or_init_from_cell_fields": (C.c_int, [vp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_ORACLE = None


def oracle():
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = _load_oracle()
    return _ORACLE


def ref_available():
    return os.path.exists(REF_SO)


_REF = None


def ref():
    """The reference's own functions, compiled from its sources (oracle/Makefile `ref`)."""
    global _REF
    if _REF is None:
        lib = C.CDLL(REF_SO)
        sig = {
            "ref_configure": (C.c_int, [C.POINTER(Params)]),
            "ref_constants": (None, [_dp]),
            "ref_set_seed": (None, [C.c_uint, C.c_uint]),
            "ref_brng": (C.c_double, []),
            "ref_irand_poisson": (C.c_int, [C.c_double]),
            "ref_isotropic": (None, [_dp]),
            "ref_initial_conditions": (None, []),
            "ref_np": (C.c_int, []),
            "ref_get_particles": (None, [_dp] * 5),
            "ref_set_particles": (C.c_int, [C.c_int] + [_dp] * 5),
            "ref_set_step": (None, [C.c_int]),
            "ref_move": (C.c_int, []),
            "ref_boundary": (None, []),
            "ref_indexing": (None, [_u32p, _u32p]),
            "ref_collide": (C.c_int, []),
            "ref_sampling": (None, []),
            "ref_clear_cells": (None, []),
            "ref_sample_step": (C.c_int, []),
            "ref_get_cells": (None, [_dp]),
            "ref_drag": (None, [_dp, C.POINTER(C.c_longlong)]),
            "ref_printing": (None, [_dp, C.c_int]),
            "ref_results_directory": (C.c_char_p, []),
            "ref_step": (C.c_int, []),
            "ref_intersect_box": (C.c_int, [C.c_double] * 8 + [_dp, _dp, _dp]),
            "ref_inside_box": (C.c_int, [C.c_double] * 6),
            "ref_diffuse": (None, [_dp] + [C.c_double] * 4),
            "ref_cll": (C.c_int, [_dp] + [C.c_double] * 6),
            "ref_generate": (C.c_int, [C.c_double] * 4 + [C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _REF = lib
    return _REF


class Draws:
    """A draw source of the oracle: MWC, Philox (k0,k1,c0,c1,c2) or a fed array."""

    def __init__(self, kind, *args):
        lib = oracle()
        self._keep = None
        if kind == "mwc":
            self.h = lib.or_draws_new_mwc(*args)
        elif kind == "philox":
            self.h = lib.or_draws_new_philox(*args)
        else:
            arr = np.ascontiguousarray(args[0], dtype=np.float64)
            self._keep = arr
            self.h = lib.or_draws_new_array(_ptr(arr), arr.size)

    def next(self):
        return oracle().or_draws_next(self.h)

    @property
    def used(self):
        return oracle().or_draws_used(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            oracle().or_draws_free(self.h)
            self.h = None


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    oracle().or_philox4x32_10(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(out, _u32p))
    return out


def constants(params):
    k = Constants()
    rc = oracle().or_constants_of(C.byref(params), C.byref(k))
    return rc, k


class Oracle:
    """One case of the oracle in REF or GPU mode."""

    def __init__(self, params, mode=MODE_GPU):
        lib = oracle()
        st = C.c_int(0)
        self.params = params
        self.h = lib.or_create(C.byref(params), mode, C.byref(st))
        if not self.h:
            raise ValueError(f"oracle create failed with code {st.value}")
        self.mode = mode
        self.ncells = params.nx * params.ny

    def __del__(self):
        if getattr(self, "h", None):
            oracle().or_destroy(self.h)
            self.h = None

    def set_mwc_seed(self, u, v):
        oracle().or_set_mwc_seed(self.h, u, v)

    def init_freestream(self):
        return oracle().or_init_freestream(self.h)

    def set_particles(self, x, y, vx, vy, vz):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, vx, vy, vz)]
        return oracle().or_set_particles(self.h, arrs[0].size, *[_ptr(a) for a in arrs])

    @property
    def np(self):
        return oracle().or_num_particles(self.h)

    def particles(self):
        n = self.np
        out = [np.zeros(n) for _ in range(5)]
        oracle().or_get_particles(self.h, *[_ptr(a) for a in out])
        return out

    def set_step(self, step):
        oracle().or_set_step(self.h, step)

    def move(self):
        return oracle().or_move(self.h)

    def boundary(self):
        return oracle().or_boundary(self.h)

    def strip_counts(self):
        out = np.zeros(4, dtype=np.int64)
        oracle().or_last_strip_counts(self.h, _ptr(out, _i64p))
        return out

    def index(self):
        return oracle().or_index(self.h)

    def get_index(self):
        counts = np.zeros(self.ncells, dtype=np.uint32)
        perm = np.zeros(self.np, dtype=np.uint32)
        oracle().or_get_index(self.h, _ptr(counts, _u32p), _ptr(perm, _u32p))
        return counts, perm

    def collide(self):
        nc = C.c_int64(0)
        rc = oracle().or_collide(self.h, C.byref(nc))
        if rc:
            raise RuntimeError(f"or_collide {rc}")
        return nc.value

    def sample(self):
        return oracle().or_sample(self.h)

    def step(self, n=1):
        return oracle().or_step(self.h, n)

    def cells(self):
        out = np.zeros((self.ncells, 17))
        oracle().or_get_cells(self.h, _ptr(out))
        return out

    def fields(self):
        out = np.zeros((self.ncells, 6))
        oracle().or_get_fields(self.h, _ptr(out))
        return out

    def state(self):
        s = State()
        oracle().or_get_state(self.h, C.byref(s))
        return s

    def init_from_cell_fields(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        return oracle().or_init_from_cell_fields(self.h, _ptr(f))


def ref_particles():
    r = ref()
    n = r.ref_np()
    out = [np.zeros(n) for _ in range(5)]
    r.ref_get_particles(*[_ptr(a) for a in out])
    return out


def ref_set_particles(x, y, vx, vy, vz):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, vx, vy, vz)]
    return ref().ref_set_particles(arrs[0].size, *[_ptr(a) for a in arrs])


def ref_cells(ncells):
    out = np.zeros((ncells, 17))
    ref().ref_get_cells(_ptr(out))
    return out


def ref_index(ncells, n):
    counts = np.zeros(ncells, dtype=np.uint32)
    perm = np.zeros(n, dtype=np.uint32)
    ref().ref_indexing(_ptr(counts, _u32p), _ptr(perm, _u32p))
    return counts, perm
