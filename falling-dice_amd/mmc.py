"""Python binding of the gfx950 beam-lattice compression solver (libmmc.so).

ctypes mirror of ``include/mmc.h``: the configuration/lattice/result structs,
the reference defaults of ``MC/include/sim.hpp:16-26`` (MC =
projects/metamaterial-compression/industry-grade), the simple-cubic workload
presets, and thin wrappers over the C-ABI entry points.  The wrappers raise
``MMCError`` with the library's own message on any non-zero status, the way the
reference's Simulator raises ``std::runtime_error`` (MC/src/sim.cpp:212,412-418).

The library is loaded from this directory; there is no CPU fallback: importing
the GPU entry points without a built ``libmmc.so`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmmc.so"

MMC_FP64, MMC_MIXED = 0, 1
MMC_OK, MMC_ERR_ARG, MMC_ERR_HIP, MMC_ERR_STATE, MMC_ERR_EQUILIBRIUM, MMC_ERR_CAPACITY = 0, 1, 2, 3, 4, 5
MMC_LATTICE_BCC, MMC_LATTICE_CUBIC = 0, 1
TRIAL_STRIDE = 17


class Material(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "density", "young", "nu", "yield_", "hardening", "bend_hardening", "torsion_hardening")]


class Config(C.Structure):
    """mmc_config — SimulationConfig (MC/include/sim.hpp:17-26) plus additive fields."""
    _fields_ = [
        ("cells_x", C.c_int32), ("cells_y", C.c_int32), ("cells_z", C.c_int32),
        ("elements_per_strut", C.c_int32),
        ("cell_size", C.c_double), ("min_radius", C.c_double), ("max_radius", C.c_double),
        ("joint_scale", C.c_double), ("trigger_scale", C.c_double),
        ("imperfection", C.c_double), ("member_imperfection", C.c_double),
        ("material", Material),
        ("gravity", C.c_double), ("shear_correction", C.c_double),
        ("contact_stiffness", C.c_double), ("contact_margin", C.c_double),
        ("centering_stiffness", C.c_double),
        ("contact_samples", C.c_int32), ("load_steps", C.c_int32), ("unload_steps", C.c_int32),
        ("max_iterations", C.c_int32),
        ("history", C.c_int32), ("max_line_search", C.c_int32),
        ("max_adaptive_refinements", C.c_int32), ("fps", C.c_int32),
        ("frames", C.c_int32), ("threads", C.c_int32),
        ("max_strain", C.c_double), ("step_time", C.c_double), ("translation_scale", C.c_double),
        ("max_translation_step", C.c_double), ("max_rotation_step", C.c_double),
        ("gradient_rms_tolerance", C.c_double), ("gradient_max_tolerance", C.c_double),
        ("armijo", C.c_double), ("min_line_step", C.c_double),
        ("penetration_tolerance", C.c_double),
        ("seed", C.c_uint64),
        ("lattice", C.c_int32), ("relative_escape", C.c_int32),
        ("relative_tolerance", C.c_double), ("position_sigma", C.c_double),
        ("radius_mode", C.c_int32), ("line_search", C.c_int32),
    ]

    def replace(self, **kw) -> "Config":
        c = Config()
        C.memmove(C.byref(c), C.byref(self), C.sizeof(Config))
        for k, v in kw.items():
            if k in ("young", "nu", "yield_", "hardening", "density", "bend_hardening",
                     "torsion_hardening"):
                setattr(c.material, k, v)
            else:
                setattr(c, k, v)
        return c


class Lattice(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("n_beams", C.c_int32),
        ("x", C.POINTER(C.c_double)), ("q", C.POINTER(C.c_double)),
        ("mass", C.POINTER(C.c_double)), ("locked", C.POINTER(C.c_uint8)),
        ("beam_a", C.POINTER(C.c_int32)), ("beam_b", C.POINTER(C.c_int32)),
        ("L0", C.POINTER(C.c_double)), ("radius", C.POINTER(C.c_double)),
        ("area", C.POINTER(C.c_double)), ("inertia", C.POINTER(C.c_double)),
        ("e0", C.POINTER(C.c_double)), ("plastic", C.POINTER(C.c_double)),
        ("n_connected", C.c_int64), ("connected", C.POINTER(C.c_uint64)),
        ("node_layer", C.POINTER(C.c_int32)), ("beam_layer", C.POINTER(C.c_int32)),
        ("beam_family", C.POINTER(C.c_int32)),
        ("width", C.c_double), ("height", C.c_double), ("depth", C.c_double),
    ]


class ObjectiveOut(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "energy", "elastic", "plastic", "contact", "reaction", "max_pen", "rms", "max")] + [
        ("contacts", C.c_int32), ("finite", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OptStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("backtracks", C.c_int32), ("rms", C.c_double),
                ("max", C.c_double), ("energy", C.c_double), ("converged", C.c_int32),
                ("evaluations", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


FRAME_FIELDS = [
    "time", "platen_y", "displacement", "engineering_strain", "engineering_stress",
    "reaction_force", "loading_work", "recovered_work", "dissipated_work",
    "specific_energy_absorption", "elastic_energy", "plastic_dissipation", "contact_energy",
    "residual_rms", "residual_max", "max_penetration", "mean_plastic_strain",
    "max_plastic_strain", "mean_plastic_rotation", "max_plastic_rotation", "permanent_set"]


class Frame(C.Structure):
    """mmc_frame — byte layout of DiskDiagnostics (MC/src/main.cpp:24-31)."""
    _fields_ = [(n, C.c_double) for n in FRAME_FIELDS] + [
        ("active_contacts", C.c_int32), ("plastic_beams", C.c_int32),
        ("plastic_hinges", C.c_int32), ("iterations", C.c_int32),
        ("converged", C.c_uint32), ("loading", C.c_uint32)]


FRAME_DTYPE = np.dtype([(n, "<f8") for n in FRAME_FIELDS] + [
    ("active_contacts", "<i4"), ("plastic_beams", "<i4"), ("plastic_hinges", "<i4"),
    ("iterations", "<i4"), ("converged", "<u4"), ("loading", "<u4")])
assert FRAME_DTYPE.itemsize == C.sizeof(Frame) == 192

# One captured body (mmc_snapshot; DiskSnapshot of the reference's META3 cache)
SNAPSHOT_DTYPE = np.dtype([("p", "<f8", (3,)), ("q", "<f8", (4,)), ("v", "<f8", (3,)), ("w", "<f8", (3,)),
                           ("half", "<f8", (3,)), ("material", "<i4"), ("id", "<i4"), ("kind", "<i4"),
                           ("module_class", "<i4"), ("ring", "<i4"), ("slot", "<i4")])
assert SNAPSHOT_DTYPE.itemsize == 152
MACHINE_BODIES = 7
# mmc_batch_frame_fn(lattice, frame, bodies, count, user)
FRAME_OBSERVER = C.CFUNCTYPE(None, C.c_int32, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p)


class LatticeInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_beams", C.c_int32), ("n_exclusions", C.c_int64),
                ("width", C.c_double), ("height", C.c_double), ("depth", C.c_double), ("mass", C.c_double)]


class StlReport(C.Structure):
    _fields_ = [("voxel", C.c_double), ("signed_volume", C.c_double), ("nx", C.c_int32), ("ny", C.c_int32),
                ("nz", C.c_int32), ("watertight", C.c_int32)] + [
        (n, C.c_int64) for n in ("triangles", "unique_edges", "initial_boundary_edges", "repaired_loops",
                                 "boundary_edges", "nonmanifold_edges")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class StateStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "sum_axial", "max_axial", "sum_rot", "max_rot", "y_min", "y_max")] + [
        ("plastic_beams", C.c_int32), ("plastic_hinges", C.c_int32)]


def default_config() -> Config:
    """Reference defaults, MC/include/sim.hpp:16-26 (BCC lattice)."""
    c = Config()
    c.cells_x, c.cells_y, c.cells_z = 4, 6, 4
    c.cell_size, c.min_radius, c.max_radius = 0.010, 0.00082, 0.00122
    c.joint_scale, c.trigger_scale = 1.18, 0.38
    c.imperfection = c.member_imperfection = 0.00022
    c.elements_per_strut = 3
    m = c.material
    m.density, m.young, m.nu, m.yield_, m.hardening = 1180.0, 26e6, 0.485, 5.2e6, 1.4e6
    m.bend_hardening, m.torsion_hardening = 0.06, 0.045
    c.gravity, c.shear_correction, c.contact_stiffness = 9.81, 0.62, 1.2e6
    c.contact_margin, c.centering_stiffness = 2e-5, 1.5
    c.contact_samples, c.load_steps, c.unload_steps = 3, 52, 30
    c.max_iterations, c.history, c.max_line_search = 130, 10, 22
    c.max_strain, c.step_time, c.translation_scale = 0.62, 0.06, 0.010
    c.max_translation_step, c.max_rotation_step = 0.0022, 0.30
    c.gradient_rms_tolerance, c.gradient_max_tolerance = 1.2e-3, 1.5e-2
    c.armijo, c.min_line_step, c.penetration_tolerance = 1e-4, 1e-9, 0.00012
    c.max_adaptive_refinements = 7
    c.fps, c.frames, c.threads = 24, 83, 0
    c.seed = 0x494E445553545259
    c.lattice = MMC_LATTICE_BCC
    c.relative_escape, c.relative_tolerance = 1, 1.5e-3
    c.position_sigma, c.radius_mode = 0.0, 0
    return c


def quick_config() -> Config:
    """The reference's --quick preset (MC/src/main.cpp:176)."""
    return default_config().replace(cells_x=3, cells_y=4, cells_z=3, load_steps=24,
                                    unload_steps=14, max_iterations=95)


def cubic_config(nx: int, ny: int, nz: int, *, cell=0.005, r_bot=0.0004, r_top=0.0006,
                 strain=0.30, load=60, unload=60, grad_tol=1e-6, max_iterations=4000,
                 sigma=0.0, radius_mode=0) -> Config:
    """Simple-cubic workload family: E 2.0 GPa, nu 0.35, yield 50 MPa, linear
    hardening 5% E, L-BFGS m = 10, absolute gradient tolerance grad_tol with the
    relative escape off.  Imperfection: the reference's corner-node pattern
    scaled to the cell (2.2e-4 m per 10 mm cell) unless sigma > 0."""
    c = default_config().replace(
        cells_x=nx, cells_y=ny, cells_z=nz, cell_size=cell, min_radius=r_bot,
        max_radius=r_top, imperfection=2.2e-4 * cell / 0.010, member_imperfection=0.0,
        young=2.0e9, nu=0.35, yield_=50e6, hardening=0.05 * 2.0e9,
        load_steps=load, unload_steps=unload, max_strain=strain,
        max_iterations=max_iterations, history=10,
        gradient_rms_tolerance=grad_tol, gradient_max_tolerance=grad_tol,
        lattice=MMC_LATTICE_CUBIC, relative_escape=0, position_sigma=sigma,
        radius_mode=radius_mode)
    c.translation_scale = cell
    return c


def specimen_config() -> Config:
    """The reference's 2x4x2 graded BCC specimen (MC/docs/VALIDATION.md:9-11)
    under its coarse convergence schedule: 10 + 7 increments to 50% strain,
    500 iterations, contact stiffness 1e6 (MC/src/main.cpp:213-221)."""
    return default_config().replace(cells_x=2, cells_y=4, cells_z=2, max_strain=0.5, load_steps=10,
                                    unload_steps=7, max_iterations=500, contact_stiffness=1e6)


def tight(cfg: Config, tol: float = 1e-7) -> Config:
    """The parity form of a workload: gradient tolerance tol (rms and max)
    with the per-term Armijo test and its approximate-Wolfe completion below the
    energy's resolution (line_search 2), relative escape off."""
    return cfg.replace(line_search=2, gradient_rms_tolerance=tol, gradient_max_tolerance=tol,
                       relative_escape=0)


def splitmix64(v: int) -> int:
    v = (v + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return v ^ (v >> 31)


def batch_workload(kind: str, count: int, load: int = 60, unload: int = 60, indices=None):
    """Seeded lattice sets for the batched solver.
      uneven: cubes of n^3 cells, n uniform in 4..16, linear radius grading with
              endpoints uniform in 0.3..0.7 mm, 30% strain
      sweep:  6^3 cells, E log-uniform 1..3 GPa, yield uniform 30..80 MPa,
              hardening uniform 2..10% of E, 40% strain
    Lattice i draws from its own stream seeded with splitmix64(seed ^ i), so a
    subset (indices) has the same lattices as the whole set."""
    import numpy as np
    seed = 0x494E445553545259
    cfgs = []
    for i in (range(count) if indices is None else indices):
        rng = np.random.default_rng(splitmix64(seed ^ i))
        if kind == "uneven":
            n = int(rng.integers(4, 17))
            r0, r1 = rng.uniform(3e-4, 7e-4, 2)
            c = cubic_config(n, n, n, r_bot=float(r0), r_top=float(r1), strain=0.30, load=load, unload=unload,
                             max_iterations=20000)
        elif kind == "sweep":
            E = float(np.exp(rng.uniform(np.log(1e9), np.log(3e9))))
            c = cubic_config(6, 6, 6, strain=0.40, load=load, unload=unload, max_iterations=20000)
            c = c.replace(young=E, yield_=float(rng.uniform(30e6, 80e6)), hardening=float(rng.uniform(0.02, 0.10)) * E)
        else:
            raise ValueError(kind)
        cfgs.append(c)
    return cfgs


def large_config(nx=64, ny=64, nz=32, *, grad_tol=1e-6, max_iterations=20000, load=100, unload=100) -> Config:
    """Contact-heavy workload: cubic cells of 5 mm, node positions jittered by a
    seeded Gaussian of 2% of the cell, per-beam radii uniform 0.3..0.7 mm,
    platen to 50% strain in 100 + 100 increments."""
    cell = 0.005
    return cubic_config(nx, ny, nz, cell=cell, r_bot=3e-4, r_top=7e-4, strain=0.5, load=load, unload=unload,
                        grad_tol=grad_tol, max_iterations=max_iterations, sigma=0.02 * cell, radius_mode=1)


# ---------------------------------------------------------------------------
# GPU library
# ---------------------------------------------------------------------------
class MMCError(RuntimeError):
    pass


_lib = None


def lib():
    """Loads libmmc.so (built by ``__graft_entry__.build()``).  Raises if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise MMCError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.mmc_create.argtypes = [C.POINTER(Lattice), C.POINTER(Config), C.c_int32, C.POINTER(P)]
        L.mmc_destroy.argtypes = [P]
        L.mmc_destroy.restype = None
        L.mmc_last_error.argtypes = [P]
        L.mmc_last_error.restype = C.c_char_p
        L.mmc_objective.argtypes = [P, dp, C.c_double, C.c_int32, C.POINTER(ObjectiveOut), dp]
        L.mmc_minimize.argtypes = [P, C.c_double, C.POINTER(OptStats), C.POINTER(ObjectiveOut), dp]
        L.mmc_commit.argtypes = [P]
        L.mmc_read_state.argtypes = [P, dp, dp, dp]
        L.mmc_read_trial.argtypes = [P, dp]
        L.mmc_read_stats.argtypes = [P, C.POINTER(StateStats)]
        L.mmc_beam_potential.argtypes = [P, dp, dp, dp]
        L.mmc_beam_potential_noalloc.argtypes = [P, P, P, P, P]
        L.mmc_capture.argtypes = [P, dp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), P]
        L.mmc_create_generated.argtypes = [C.POINTER(Config), C.c_int32, C.c_int32, C.POINTER(P)]
        L.mmc_lattice_info_get.argtypes = [P, C.c_int32, C.POINTER(LatticeInfo)]
        i32p = C.POINTER(C.c_int32)
        L.mmc_read_lattice.argtypes = [P, C.c_int32, dp, dp, C.POINTER(C.c_uint8), i32p, i32p, dp, dp, dp, dp, dp,
                                       i32p, i32p, i32p, i32p, i32p]
        L.mmc_batch_create_generated.argtypes = [C.POINTER(Config), C.c_int32, C.c_int32, C.POINTER(P)]
        L.mmc_capture_mark.argtypes = [P]
        L.mmc_config_defaults.argtypes = [C.POINTER(Config)]
        L.mmc_config_defaults.restype = None
        L.mmc_sim_create.argtypes = [C.POINTER(Config), C.c_int32, C.POINTER(P)]
        L.mmc_sim_destroy.argtypes = [P]
        L.mmc_sim_destroy.restype = None
        L.mmc_sim_last_error.argtypes = [P]
        L.mmc_sim_last_error.restype = C.c_char_p
        L.mmc_sim_info.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), dp, dp, dp]
        L.mmc_sim_lattice.argtypes = [P, dp, dp, C.POINTER(C.c_uint8), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32), dp, dp, dp]
        L.mmc_sim_run.argtypes = [P, C.POINTER(Frame), C.c_int32, C.POINTER(C.c_int32), C.c_int32]
        L.mmc_sim_rigid_rotation_test.argtypes = [P, dp, C.POINTER(C.c_int32)]
        L.mmc_sim_gradient_check.argtypes = [P, dp, C.POINTER(C.c_int32)]
        L.mmc_sim_context.argtypes = [P]
        L.mmc_sim_context.restype = P
        L.mmc_export_stl.argtypes = [C.POINTER(Lattice), C.c_double, C.c_int32, C.c_char_p, C.POINTER(StlReport)]
        L.mmc_stl_last_error.argtypes = []
        L.mmc_stl_last_error.restype = C.c_char_p
        L.mmc_stl_case_table.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int8)]
        L.mmc_stl_case_table.restype = None
        L.mmc_sim_export_stl.argtypes = [P, C.c_char_p, C.c_int32, C.POINTER(StlReport)]
        L.mmc_sim_begin.argtypes = [P]
        L.mmc_sim_advance.argtypes = [P, C.POINTER(Frame), C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]
        L.mmc_kernel_timer.argtypes = [P, C.c_int32]
        L.mmc_kernel_timer_read.argtypes = [P, C.POINTER(C.c_int64), dp]
        L.mmc_time_kernel.argtypes = [P, C.c_int32, C.c_int32, dp]
        L.mmc_loop_shape.argtypes = [P, C.POINTER(C.c_int32), dp]
        L.mmc_evaluations.argtypes = [P]
        L.mmc_evaluations.restype = C.c_int64
        L.mmc_batch_create.argtypes = [C.POINTER(Config), C.c_int32, C.c_int32, C.POINTER(P)]
        L.mmc_batch_destroy.argtypes = [P]
        L.mmc_batch_destroy.restype = None
        L.mmc_batch_last_error.argtypes = [P]
        L.mmc_batch_last_error.restype = C.c_char_p
        L.mmc_batch_info.argtypes = [P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), dp, dp, dp]
        L.mmc_batch_run.argtypes = [P, C.POINTER(Frame), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.c_int32]
        L.mmc_batch_set_frame_observer.argtypes = [P, FRAME_OBSERVER, P]
        L.mmc_batch_evaluations.argtypes = [P]
        L.mmc_batch_evaluations.restype = C.c_int64
        L.mmc_batch_context.argtypes = [P]
        L.mmc_batch_context.restype = P
        L.mmc_lattice_build.argtypes = [C.POINTER(Config), C.POINTER(Lattice), C.POINTER(P)]
        L.mmc_lattice_free.argtypes = [P]
        L.mmc_lattice_free.restype = None
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class LatticeArrays:
    """Owns numpy arrays for an mmc_lattice (node-major AoS)."""

    def __init__(self, **arrays):
        self.__dict__.update(arrays)

    def struct(self) -> Lattice:
        L = Lattice()
        L.n_nodes, L.n_beams = len(self.mass), len(self.L0)
        L.x, L.q, L.mass = _dp(self.x), _dp(self.q), _dp(self.mass)
        L.locked = self.locked.ctypes.data_as(C.POINTER(C.c_uint8))
        L.beam_a = self.a.ctypes.data_as(C.POINTER(C.c_int32))
        L.beam_b = self.b.ctypes.data_as(C.POINTER(C.c_int32))
        L.L0, L.radius, L.area, L.inertia = _dp(self.L0), _dp(self.radius), _dp(self.area), _dp(self.inertia)
        L.e0, L.plastic = _dp(self.e0), _dp(self.plastic)
        L.n_connected = len(self.connected)
        L.connected = self.connected.ctypes.data_as(C.POINTER(C.c_uint64))
        L.node_layer = self.node_layer.ctypes.data_as(C.POINTER(C.c_int32))
        L.beam_layer = self.beam_layer.ctypes.data_as(C.POINTER(C.c_int32))
        L.beam_family = self.beam_family.ctypes.data_as(C.POINTER(C.c_int32))
        L.width, L.height, L.depth = self.width, self.height, self.depth
        return L


def build_lattice(cfg: Config) -> LatticeArrays:
    """The facade's host lattice for cfg (mmc_lattice_build; no GPU work)."""
    L = lib()
    view, h = Lattice(), C.c_void_p()
    if L.mmc_lattice_build(C.byref(cfg), C.byref(view), C.byref(h)) != 0:
        raise MMCError(L.mmc_sim_last_error(None).decode())
    try:
        N, B, NC = view.n_nodes, view.n_beams, view.n_connected

        def cp(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)
        arr = LatticeArrays(
            x=cp(view.x, 3 * N, np.float64),
            q=np.tile(np.array([1.0, 0, 0, 0]), N),
            mass=cp(view.mass, N, np.float64), locked=cp(view.locked, 6 * N, np.uint8),
            a=cp(view.beam_a, B, np.int32), b=cp(view.beam_b, B, np.int32),
            L0=cp(view.L0, B, np.float64), radius=cp(view.radius, B, np.float64),
            area=cp(view.area, B, np.float64), inertia=cp(view.inertia, B, np.float64),
            e0=cp(view.e0, 3 * B, np.float64), plastic=np.zeros(7 * B),
            connected=cp(view.connected, NC, np.uint64),
            node_layer=cp(view.node_layer, N, np.int32), beam_layer=cp(view.beam_layer, B, np.int32),
            beam_family=cp(view.beam_family, B, np.int32),
            width=view.width, height=view.height, depth=view.depth)
    finally:
        L.mmc_lattice_free(h)
    return arr


class Context:
    """One device context (mmc_ctx): the hot path of a single lattice."""

    def __init__(self, lat: LatticeArrays | None, cfg: Config, precision: int = MMC_FP64):
        L = lib()
        self.cfg = cfg
        h = C.c_void_p()
        if lat is None:
            # the lattice built on the device (mmc_create_generated)
            rc = L.mmc_create_generated(C.byref(cfg), 1, precision, C.byref(h))
            if rc != 0:
                raise MMCError(f"mmc_create_generated failed ({rc}): {L.mmc_last_error(None).decode()}")
            self.h = h
            info = self.lattice_info(0)
            self.n_nodes, self.n_beams = info.n_nodes, info.n_beams
            return
        self._lat = lat
        self._ls = lat.struct()
        self.n_nodes, self.n_beams = len(lat.mass), len(lat.L0)
        rc = L.mmc_create(C.byref(self._ls), C.byref(cfg), precision, C.byref(h))
        if rc != 0:
            raise MMCError(f"mmc_create failed ({rc}): {L.mmc_last_error(None).decode()}")
        self.h = h

    @classmethod
    def generated(cls, cfg: Config, precision: int = MMC_FP64) -> "Context":
        """A context whose lattice is built on the device."""
        return cls(None, cfg, precision)

    def lattice_info(self, b: int = 0) -> LatticeInfo:
        return lattice_info(self.h, b)

    def read_lattice(self, b: int = 0):
        """Lattice b's construction as the device holds it: (LatticeArrays
        without plastic state and exclusion keys, adj_off, adj)."""
        return read_lattice(self.h, b)

    def _check(self, rc):
        if rc != 0:
            raise MMCError(lib().mmc_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().mmc_destroy(self.h)
            self.h = None

    __del__ = close

    def objective(self, z, top, trials=False, grad=True):
        z = np.ascontiguousarray(z, dtype=np.float64)
        out = ObjectiveOut()
        g = np.zeros(6 * self.n_nodes) if grad else None
        self._check(lib().mmc_objective(self.h, _dp(z), top, int(trials), C.byref(out), _dp(g)))
        return out, g

    def minimize(self, top):
        st, fin = OptStats(), ObjectiveOut()
        z = np.zeros(6 * self.n_nodes)
        self._check(lib().mmc_minimize(self.h, top, C.byref(st), C.byref(fin), _dp(z)))
        return st, fin, z

    def commit(self):
        self._check(lib().mmc_commit(self.h))

    def read_state(self):
        x = np.zeros(3 * self.n_nodes)
        q = np.zeros(4 * self.n_nodes)
        p = np.zeros(7 * self.n_beams)
        self._check(lib().mmc_read_state(self.h, _dp(x), _dp(q), _dp(p)))
        return x, q, p

    def read_trial(self):
        t = np.zeros(TRIAL_STRIDE * self.n_beams)
        self._check(lib().mmc_read_trial(self.h, _dp(t)))
        return t.reshape(self.n_beams, TRIAL_STRIDE)

    def stats(self):
        s = StateStats()
        self._check(lib().mmc_read_stats(self.h, C.byref(s)))
        return s

    def beam_potential(self, x, q):
        out = np.zeros(1)
        x = np.ascontiguousarray(x, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._check(lib().mmc_beam_potential(self.h, _dp(x), _dp(q), _dp(out)))
        return float(out[0])

    def capture_mark(self):
        """Stores the committed poses as the previous poses of the next capture."""
        self._check(lib().mmc_capture_mark(self.h))

    def capture(self, top, velocities=False):
        """The bodies of the committed frame (SNAPSHOT_DTYPE, B + N + 7), with
        velocities against the poses of the last capture_mark when asked."""
        return capture(self.h, [top], [1], [int(velocities)], self.n_beams + self.n_nodes + MACHINE_BODIES)


def lattice_info(ctx, b: int) -> LatticeInfo:
    info = LatticeInfo()
    if lib().mmc_lattice_info_get(ctx, b, C.byref(info)) != 0:
        raise MMCError(lib().mmc_last_error(ctx).decode())
    return info


def read_lattice(ctx, b: int):
    """mmc_read_lattice of lattice b of a context handle."""
    info = lattice_info(ctx, b)
    N, B = info.n_nodes, info.n_beams
    a = dict(x=np.zeros(3 * N), mass=np.zeros(N), locked=np.zeros(6 * N, np.uint8), a=np.zeros(B, np.int32),
             b=np.zeros(B, np.int32), L0=np.zeros(B), radius=np.zeros(B), area=np.zeros(B), inertia=np.zeros(B),
             e0=np.zeros(3 * B), node_layer=np.zeros(N, np.int32), beam_layer=np.zeros(B, np.int32),
             beam_family=np.zeros(B, np.int32))
    adj_off = np.zeros(N + 1, np.int32)
    adj = np.zeros(max(1, 2 * info.n_exclusions), np.int32)
    i32 = lambda v: v.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
    rc = lib().mmc_read_lattice(ctx, b, _dp(a["x"]), _dp(a["mass"]), a["locked"].ctypes.data_as(C.POINTER(C.c_uint8)),
                                i32(a["a"]), i32(a["b"]), _dp(a["L0"]), _dp(a["radius"]), _dp(a["area"]),
                                _dp(a["inertia"]), _dp(a["e0"]), i32(a["node_layer"]), i32(a["beam_layer"]),
                                i32(a["beam_family"]), i32(adj_off), i32(adj))
    if rc != 0:
        raise MMCError(lib().mmc_last_error(ctx).decode())
    lat = LatticeArrays(width=info.width, height=info.height, depth=info.depth, q=None, plastic=None,
                        connected=np.zeros(0, np.uint64), **a)
    return lat, adj_off, adj[:2 * info.n_exclusions]


def capture(ctx, tops, mask, velocities, total):
    """mmc_capture on a context handle: `total` bodies of the masked lattices."""
    out = np.zeros(total, SNAPSHOT_DTYPE)
    t = np.ascontiguousarray(tops, dtype=np.float64)
    m = (C.c_int32 * len(mask))(*mask)
    v = (C.c_int32 * len(velocities))(*velocities)
    rc = lib().mmc_capture(ctx, _dp(t), m, v, out.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise MMCError(lib().mmc_last_error(ctx).decode())
    return out


class Simulator:
    """The reference's Simulator facade (MC/include/sim.hpp:35-54) over mmc_sim_*."""

    def __init__(self, cfg: Config, precision: int = MMC_FP64):
        L = lib()
        h = C.c_void_p()
        self.cfg = cfg
        rc = L.mmc_sim_create(C.byref(cfg), precision, C.byref(h))
        if rc != 0:
            msg = L.mmc_sim_last_error(h).decode() if h else L.mmc_last_error(None).decode()
            if h:
                L.mmc_sim_destroy(h)
            raise MMCError(msg)
        self.h = h
        nn, nb = C.c_int32(), C.c_int32()
        hh, m, a = C.c_double(), C.c_double(), C.c_double()
        L.mmc_sim_info(h, C.byref(nn), C.byref(nb), C.byref(hh), C.byref(m), C.byref(a))
        self.n_nodes, self.n_beams = nn.value, nb.value
        self.height, self.mass, self.area = hh.value, m.value, a.value

    def close(self):
        if getattr(self, "h", None):
            lib().mmc_sim_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc):
        if rc != 0:
            raise MMCError(lib().mmc_sim_last_error(self.h).decode())

    def run(self, verbose=False):
        cap = 4 * (self.cfg.load_steps + self.cfg.unload_steps + 1) + 4096
        frames = (Frame * cap)()
        n = C.c_int32()
        self._check(lib().mmc_sim_run(self.h, frames, cap, C.byref(n), int(verbose)))
        return np.frombuffer(bytes(frames)[: n.value * C.sizeof(Frame)], dtype=FRAME_DTYPE).copy()

    def rigid_rotation_test(self):
        e, p = C.c_double(), C.c_int32()
        self._check(lib().mmc_sim_rigid_rotation_test(self.h, C.byref(e), C.byref(p)))
        return bool(p.value), e.value

    def gradient_check(self):
        e, p = C.c_double(), C.c_int32()
        self._check(lib().mmc_sim_gradient_check(self.h, C.byref(e), C.byref(p)))
        return bool(p.value), e.value

    def lattice(self):
        N, B = self.n_nodes, self.n_beams
        x, m = np.zeros(3 * N), np.zeros(N)
        lk = np.zeros(6 * N, np.uint8)
        a, b = np.zeros(B, np.int32), np.zeros(B, np.int32)
        L0, r, e0 = np.zeros(B), np.zeros(B), np.zeros(3 * B)
        self._check(lib().mmc_sim_lattice(
            self.h, _dp(x), _dp(m), lk.ctypes.data_as(C.POINTER(C.c_uint8)),
            a.ctypes.data_as(C.POINTER(C.c_int32)), b.ctypes.data_as(C.POINTER(C.c_int32)),
            _dp(L0), _dp(r), _dp(e0)))
        return dict(x=x, mass=m, locked=lk, a=a, b=b, L0=L0, radius=r, e0=e0)

    def context(self) -> int:
        return lib().mmc_sim_context(self.h)

    def export_stl(self, path, segments=12) -> dict:
        """Simulator::exportSTL on the device; returns the audit report."""
        rep = StlReport()
        self._check(lib().mmc_sim_export_stl(self.h, str(path).encode(), int(segments), C.byref(rep)))
        return rep.as_dict()

    def read_state(self):
        """Committed base pose (x 3N, q 4N) and plastic states (7B), node-major."""
        x, q, p = np.zeros(3 * self.n_nodes), np.zeros(4 * self.n_nodes), np.zeros(7 * self.n_beams)
        rc = lib().mmc_read_state(self.context(), _dp(x), _dp(q), _dp(p))
        if rc != 0:
            raise MMCError(lib().mmc_last_error(self.context()).decode())
        return x, q, p

    def begin(self):
        self._check(lib().mmc_sim_begin(self.h))

    def advance(self):
        """Solves the next scheduled increment; returns (frames, done)."""
        cap = 512
        frames = (Frame * cap)()
        n, done = C.c_int32(), C.c_int32()
        self._check(lib().mmc_sim_advance(self.h, frames, cap, C.byref(n), C.byref(done)))
        arr = np.frombuffer(bytes(frames)[: n.value * C.sizeof(Frame)], dtype=FRAME_DTYPE).copy()
        return arr, bool(done.value)

    def kernel_timer(self, kernel: int):
        rc = lib().mmc_kernel_timer(self.context(), kernel)
        if rc != 0:
            raise MMCError(lib().mmc_last_error(self.context()).decode())

    def kernel_timer_read(self):
        n, ms = C.c_int64(), C.c_double()
        lib().mmc_kernel_timer_read(self.context(), C.byref(n), C.byref(ms))
        return n.value, ms.value

    def evaluations(self) -> int:
        return int(lib().mmc_evaluations(self.context()))

    def loop_shape(self) -> dict:
        """Decision shape chosen by the device loop's tuner (blocks per lattice
        of the decision kernel; 0: fused into the gather kernel), and the
        gather + decision times (us) of each candidate."""
        d = C.c_int32()
        du = (C.c_double * 6)()
        lib().mmc_loop_shape(self.context(), C.byref(d), du)
        return {"decide_blocks": d.value,
                "gather_decide_us": {str(k): round(v, 3) for k, v in zip((1, 2, 4, 8, 16, "fused"), du)}}

    def time_kernel(self, kernel: str, launches: int = 200) -> float:
        """Mean duration (ms) of one objective kernel, HIP events on the solver stream."""
        ms = C.c_double()
        rc = lib().mmc_time_kernel(self.context(), KERNELS[kernel], launches, C.byref(ms))
        if rc != 0:
            raise MMCError(lib().mmc_last_error(self.context()).decode())
        return ms.value


KERNELS = {"state": 0, "fill": 2, "contact": 4, "beam": 5, "gather": 6, "final": 7, "direction": 8, "pair": 9,
           "dstate_g": 10, "elem_g": 11, "gather_g": 12, "decide_g": 13, "copy": 14}


def stl_case_table():
    """(counts[256], tris[256, 10, 3]) of the export's cell triangulation (host only)."""
    counts = np.zeros(256, np.int32)
    tris = np.zeros((256, 10, 3), np.int8)
    lib().mmc_stl_case_table(counts.ctypes.data_as(C.POINTER(C.c_int32)), tris.ctypes.data_as(C.POINTER(C.c_int8)))
    return counts, tris


def device_name() -> str:
    return os.environ.get("HIP_VISIBLE_DEVICES", "0")


class BatchSimulator:
    """Many lattices solved together (mmc_batch_*): each with the schedule of
    Simulator.run under its own config."""

    def __init__(self, cfgs, precision: int = MMC_FP64, device_lattices: bool = False):
        L = lib()
        n = len(cfgs)
        arr = (Config * n)(*cfgs)
        h = C.c_void_p()
        create = L.mmc_batch_create_generated if device_lattices else L.mmc_batch_create
        rc = create(arr, n, precision, C.byref(h))
        if rc != 0:
            msg = L.mmc_batch_last_error(h).decode() if h else L.mmc_batch_last_error(None).decode()
            if h:
                L.mmc_batch_destroy(h)
            raise MMCError(msg)
        self.h, self.n, self.cfgs = h, n, list(cfgs)
        self.n_nodes, self.n_beams = [], []
        for i in range(n):
            nn, nb = C.c_int32(), C.c_int32()
            hh, m, a = C.c_double(), C.c_double(), C.c_double()
            L.mmc_batch_info(h, i, C.byref(nn), C.byref(nb), C.byref(hh), C.byref(m), C.byref(a))
            self.n_nodes.append(nn.value)
            self.n_beams.append(nb.value)

    def close(self):
        if getattr(self, "h", None):
            lib().mmc_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def context(self) -> int:
        return lib().mmc_batch_context(self.h)

    def evaluations(self) -> int:
        return int(lib().mmc_batch_evaluations(self.h))

    def run(self, verbose=False, on_bodies=None):
        """Returns (frames per lattice, status per lattice, first error text).
        on_bodies(lattice, frame, bodies) receives every committed frame's
        bodies (SNAPSHOT_DTYPE) as the device captures them."""
        cb = None
        if on_bodies is not None:
            def observe(lattice, frame, ptr, count, _user):
                raw = C.string_at(ptr, int(count) * SNAPSHOT_DTYPE.itemsize)
                on_bodies(int(lattice), int(frame), np.frombuffer(raw, SNAPSHOT_DTYPE).copy())
            cb = FRAME_OBSERVER(observe)
            lib().mmc_batch_set_frame_observer(self.h, cb, None)
        try:
            return self._run(verbose)
        finally:
            if cb is not None:
                lib().mmc_batch_set_frame_observer(self.h, FRAME_OBSERVER(), None)

    def _run(self, verbose):
        cap = 4 * max(c.load_steps + c.unload_steps + 1 for c in self.cfgs) + 256
        frames = (Frame * (cap * self.n))()
        nf = (C.c_int32 * self.n)()
        st = (C.c_int32 * self.n)()
        rc = lib().mmc_batch_run(self.h, frames, cap, nf, st, int(verbose))
        if rc != 0:
            raise MMCError(lib().mmc_batch_last_error(self.h).decode())
        raw = np.frombuffer(bytes(frames), dtype=FRAME_DTYPE)
        out = [raw[b * cap: b * cap + min(cap, nf[b])].copy() for b in range(self.n)]
        return out, list(st), lib().mmc_batch_last_error(self.h).decode()
