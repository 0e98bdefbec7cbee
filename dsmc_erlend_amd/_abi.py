"""ctypes mirror of the POD structs and enums declared in include/dsmc_c.h."""
import ctypes as C

WALL_DIFFUSE, WALL_SPECULAR, WALL_CLL = 0, 1, 2
WALL_NAMES = {WALL_DIFFUSE: "diffuse", WALL_SPECULAR: "specular", WALL_CLL: "cll"}

OK = 0
E_CAPACITY = 1
E_ARGUMENT = 2
E_HIP = 3
E_OUTSIDE_DOMAIN = 20
E_CELL_CAPACITY = 22
E_CLL_COEFF = 40
E_IO = 50

STAGES = ["begin", "move", "extra", "scan", "scatter", "gather", "collide"]

SCHEDULE_LARGE, SCHEDULE_SMALL = 0, 1
SMALL_MAX_CELLS = 12288


class Params(C.Structure):
    _fields_ = [
        ("x1", C.c_double), ("x2", C.c_double), ("y1", C.c_double), ("y2", C.c_double),
        ("dz", C.c_double), ("nx", C.c_int32), ("ny", C.c_int32), ("dl", C.c_double),
        ("box_x", C.c_double), ("box_y", C.c_double), ("box_width", C.c_double),
        ("box_height", C.c_double),
        ("dt", C.c_double), ("n_step", C.c_int32), ("first_sampling_step", C.c_int32),
        ("sampling_period", C.c_int32),
        ("npc_free", C.c_double), ("molar_mass", C.c_double), ("mach", C.c_double),
        ("n_free", C.c_double), ("t_free", C.c_double), ("attack_angle_deg", C.c_double),
        ("t_wall", C.c_double), ("d_ref", C.c_double), ("viscosity_index", C.c_double),
        ("t_ref", C.c_double),
        ("wall_model", C.c_int32), ("acc_n", C.c_double), ("acc_t", C.c_double),
        ("seed", C.c_uint64), ("case_id", C.c_uint32),
        ("max_per_cell", C.c_uint32), ("capacity_factor", C.c_double),
    ]

    def replace(self, **kw):
        p = Params()
        C.pointer(p)[0] = self
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Constants(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "molecule_mass", "p_free", "u_free", "ux_free", "uy_free",
        "dx", "dy", "dv", "weight",
        "sigma_ref", "omega", "cr_ref", "sc_max",
        "mft_free", "mfp_free", "kn_free")]


class State(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("np", C.c_int64), ("sample_step", C.c_int64),
        ("last_collisions", C.c_int64), ("total_collisions", C.c_int64),
        ("last_injected", C.c_int64), ("last_removed", C.c_int64), ("wall_hits", C.c_int64),
        ("drag_impulse_total", C.c_double), ("drag_sampling_time", C.c_double),
        ("time", C.c_double), ("sample_time", C.c_double), ("particle_steps", C.c_int64),
        ("mean_drag", C.c_double), ("drag_coefficient", C.c_double),
    ]


class Tuning(C.Structure):
    _fields_ = [("schedule", C.c_int32), ("chunk", C.c_int32),
                ("particles_per_thread", C.c_int32)]


def default_params(**kw):
    """The reference's compile-time defaults (src/Geometry.cpp:11-24,
    src/Initialization.cpp:11-31, src/Boundary.cpp:10-13), overridden by keywords."""
    p = Params(
        x1=-20.0, x2=20.0, y1=-20.0, y2=20.0, dz=0.1, nx=100, ny=100, dl=0.1,
        box_x=0.0, box_y=0.0, box_width=4.0, box_height=4.0,
        dt=1e-5, n_step=100000, first_sampling_step=10000, sampling_period=1,
        npc_free=10.0, molar_mass=0.040, mach=4.0, n_free=1e17, t_free=200.0,
        attack_angle_deg=0.0, t_wall=300.0, d_ref=4.17e-10, viscosity_index=0.81,
        t_ref=273.0, wall_model=WALL_CLL, acc_n=0.8, acc_t=0.8, seed=1, case_id=0,
        max_per_cell=0, capacity_factor=0.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p
