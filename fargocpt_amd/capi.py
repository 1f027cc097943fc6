"""ctypes mirror of include/fcpt.h and a thin Python handle over libfcpt_hip.so.

The structures below must stay field-for-field identical to include/fcpt.h.
Loading the library fails loudly when the HIP build is missing: there is no
CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfcpt_hip.so")

FIELDS = [
    "SIGMA", "VRAD", "VAZI", "ENERGY", "PRESSURE", "SOUNDSPEED", "SCALE_HEIGHT",
    "VISCOSITY", "POTENTIAL", "Q_R", "Q_PHI", "DIV_V", "TAU_R_R", "TAU_R_PHI",
    "TAU_PHI_PHI", "QPLUS", "QMINUS", "TEMPERATURE", "SIGMA0", "VRAD0", "VAZI0", "ENERGY0",
]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
VECTOR_FIELDS = {"VRAD", "VRAD0", "TAU_R_PHI"}

ARTVISC = {"none": 0, "tw": 1, "sn": 2}
# kernels of the (fused) step that fcpt_time_kernel can time on their own
KERNELS = {"radial_sweep": 0, "azimuthal_sweep": 1, "cfl": 3, "visc_fused": 7, "potential": 5}
# the same on a batch (the viscous pass is not timed there)
BATCH_KERNELS = {"radial_sweep": 0, "azimuthal_sweep": 1, "cfl": 3, "potential": 5}
KERNEL_STRESS = 4      # the separate stress-tensor kernel of fcpt_viscous_stress_tensor
KERNEL_DISK_FORCE = 6  # the disk-on-planet reduction (not part of the hydro step)
# storage modes (fcpt_disk_desc.precision) and launch parameters of include/fcpt.h
PREC_FP64, PREC_SPLIT32 = 0, 1
INTEGRATOR_EULER, INTEGRATOR_LEAPFROG = 0, 1
PARAM_RADIAL_TILE, PARAM_RADIAL_WAVES, PARAM_FUSED_STEP, PARAM_VISC_TILE, PARAM_KICK_TILE = 0, 1, 2, 3, 4
PARAM_SPLIT_ROUNDING, PARAM_REBASE_INTERVAL = 5, 6
BC = {"reflecting": 0, "zerogradient": 1, "outflow": 2}
MAX_PLANETS = 8


class DiskDesc(C.Structure):
    _fields_ = [
        ("nr", C.c_int32), ("nphi", C.c_int32), ("radii", C.POINTER(C.c_double)),
        ("rmin", C.c_double), ("rmax", C.c_double),
        ("adiabatic", C.c_int32), ("precision", C.c_int32),
        ("G", C.c_double), ("R", C.c_double), ("sigma_sb", C.c_double), ("c_light", C.c_double),
        ("hydro_center_mass", C.c_double), ("mu", C.c_double), ("gamma", C.c_double),
        ("aspect_ratio", C.c_double), ("flaring_index", C.c_double), ("alpha", C.c_double),
        ("thickness_smoothing", C.c_double),
        ("cfl", C.c_double), ("cfl_max_var", C.c_double), ("heating_cooling_cfl_limit", C.c_double),
        ("sigma0", C.c_double), ("sigma_floor", C.c_double), ("t_min", C.c_double),
        ("t_max", C.c_double),
        ("artificial_viscosity", C.c_int32), ("artvisc_dissipation", C.c_int32),
        ("artvisc_factor", C.c_double),
        ("heating_viscous", C.c_int32), ("heating_viscous_factor", C.c_double),
        ("cooling_beta_enabled", C.c_int32), ("cooling_beta_ref", C.c_int32),
        ("cooling_beta", C.c_double), ("cooling_beta_ramp_up", C.c_double),
        ("radial_viscosity_factor", C.c_double),
        ("bc_inner", C.c_int32), ("bc_outer", C.c_int32), ("damping", C.c_int32),
        ("damping_inner_limit", C.c_double), ("damping_outer_limit", C.c_double),
        ("damping_time_factor", C.c_double), ("damping_time_radius_outer", C.c_double),
        ("damp_inner_scalar", C.c_int32), ("damp_inner_vector", C.c_int32),
        ("damp_outer_scalar", C.c_int32), ("damp_outer_vector", C.c_int32),
        ("omega_frame", C.c_double), ("monitor_timestep", C.c_double),
        ("integrator", C.c_int32), ("potential_smoothing", C.c_int32),
    ]


class Planet(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("mass", C.c_double),
                ("r_cubic", C.c_double)]


class StepParams(C.Structure):
    _fields_ = [("n_planets", C.c_int32), ("pad_", C.c_int32),
                ("planet", Planet * MAX_PLANETS),
                ("indirect_x", C.c_double), ("indirect_y", C.c_double)]


class IcDesc(C.Structure):
    """fcpt_ic_desc: the profile initial condition's own configuration keys."""
    _fields_ = [("sigma_slope", C.c_double), ("imposed_disk_drift", C.c_double),
                ("vradial_zero", C.c_int32), ("cutoff_outer", C.c_int32),
                ("cutoff_inner", C.c_int32), ("pad_", C.c_int32),
                ("cutoff_outer_point", C.c_double), ("cutoff_outer_width", C.c_double),
                ("cutoff_inner_point", C.c_double), ("cutoff_inner_width", C.c_double)]


def make_ic(sigma_slope, imposed_disk_drift=0.0, vradial_zero=0, cutoff_outer=0,
            cutoff_outer_point=1e300, cutoff_outer_width=1.0, cutoff_inner=0,
            cutoff_inner_point=0.0, cutoff_inner_width=1.0) -> IcDesc:
    return IcDesc(sigma_slope, imposed_disk_drift, vradial_zero, cutoff_outer, cutoff_inner, 0,
                  cutoff_outer_point, cutoff_outer_width, cutoff_inner_point, cutoff_inner_width)


class SnapshotInfo(C.Structure):
    """fcpt_snapshot_info: the misc.bin record of a snapshot."""
    _fields_ = [("snapshot", C.c_uint32), ("n_monitor", C.c_uint32), ("time", C.c_double),
                ("omega_frame", C.c_double), ("frame_angle", C.c_double),
                ("last_dt", C.c_double), ("n_iter", C.c_uint64)]


class ForceBody(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("r", C.c_double), ("r_cubic", C.c_double)]


FORCE_NO_STAR_SMOOTHING, FORCE_SMOOTHING_PLANETLOC, FORCE_CORRECT_SELFGRAVITY = 1, 2, 4


def make_force_bodies(bodies):
    """bodies: sequence of (x, y, r, r_cubic) -> ctypes array of ForceBody"""
    arr = (ForceBody * max(1, len(bodies)))()
    for k, (x, y, r, rc) in enumerate(bodies):
        arr[k] = ForceBody(x, y, r, rc)
    return arr


def make_step_params(planets, indirect=(0.0, 0.0)) -> StepParams:
    """planets: iterable of (x, y, mass, r_cubic); body 0 is the central object."""
    sp = StepParams()
    planets = list(planets)
    if len(planets) > MAX_PLANETS:
        raise ValueError("too many planets")
    sp.n_planets = len(planets)
    for k, (x, y, m, rc) in enumerate(planets):
        sp.planet[k] = Planet(float(x), float(y), float(m), float(rc))
    sp.indirect_x, sp.indirect_y = float(indirect[0]), float(indirect[1])
    return sp


def desc_from_dict(cfg: dict) -> tuple[DiskDesc, np.ndarray]:
    """Build a DiskDesc from a flat parameter dict (keys as in DiskDesc).

    ``cfg['radii']`` must hold nr+2 interface radii.  Returns the struct and the
    radii array, which the caller must keep alive while the struct is used.
    """
    radii = np.ascontiguousarray(cfg["radii"], dtype=np.float64)
    d = DiskDesc()
    defaults = {
        "precision": 0, "heating_cooling_cfl_limit": 1.0, "radial_viscosity_factor": 1.0,
        "cooling_beta_ramp_up": 0.0, "bc_inner": 0, "bc_outer": 0,
        "damp_inner_scalar": -1, "damp_inner_vector": -1, "damp_outer_scalar": -1,
        "damp_outer_vector": -1, "integrator": 0, "potential_smoothing": 0,
    }
    for name, _ in DiskDesc._fields_:
        if name == "radii":
            continue
        if name in cfg:
            setattr(d, name, cfg[name])
        elif name in defaults:
            setattr(d, name, defaults[name])
        else:
            raise KeyError(f"disk description lacks {name!r}")
    if len(radii) != d.nr + 2:
        raise ValueError("radii must hold nr+2 values")
    d.radii = radii.ctypes.data_as(C.POINTER(C.c_double))
    return d, radii


def field_shape(name: str, nr: int, nphi: int) -> tuple[int, int]:
    return (nr + 1, nphi) if name in VECTOR_FIELDS else (nr, nphi)


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libfcpt_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.c_void_p
    D = C.c_double
    I = C.c_int32
    dp = C.POINTER(C.c_double)
    sig = {
        "fcpt_create": [C.POINTER(DiskDesc), C.POINTER(P)],
        "fcpt_destroy": [P],
        "fcpt_synchronize": [P],
        "fcpt_upload_state": [P, dp, dp, dp, dp],
        "fcpt_download_state": [P, dp, dp, dp, dp],
        "fcpt_upload_reference": [P, dp, dp, dp, dp],
        "fcpt_upload_field": [P, I, dp],
        "fcpt_download_field": [P, I, dp],
        "fcpt_set_time": [P, D, D, I],
        "fcpt_get_time": [P, dp, dp, C.POINTER(I)],
        "fcpt_set_step_params": [P, C.POINTER(StepParams)],
        "fcpt_set_midstep_params": [P, C.POINTER(StepParams)],
        "fcpt_init_profiles": [P, C.POINTER(IcDesc)],
        "fcpt_write_snapshot": [P, C.c_char_p, C.POINTER(SnapshotInfo), C.c_uint32],
        "fcpt_read_snapshot": [P, C.c_char_p, C.POINTER(SnapshotInfo)],
        "fcpt_read_snapshot_reference": [P, C.c_char_p],
        "fcpt_write_grid_files": [P, C.c_char_p, C.c_char_p],
        "fcpt_compute_pressure": [P],
        "fcpt_step_leapfrog": [P, D],
        "fcpt_condition_cfl": [P, dp],
        "fcpt_calculate_timestep": [P, dp],
        "fcpt_nbody_potential": [P],
        "fcpt_update_with_sourceterms": [P, D],
        "fcpt_artificial_viscosity": [P, D],
        "fcpt_recalculate_viscosity": [P],
        "fcpt_viscous_stress_tensor": [P],
        "fcpt_update_velocities_with_viscosity": [P, D],
        "fcpt_substep3": [P, D, D],
        "fcpt_apply_boundary_condition": [P, D, I],
        "fcpt_transport": [P, D],
        "fcpt_recalculate_derived": [P],
        "fcpt_get_nshift": [P, C.POINTER(I)],
        "fcpt_step_euler": [P, D],
        "fcpt_run_step": [P, dp],
        "fcpt_totals": [P, D, dp, dp],
        "fcpt_get_ring_geometry": [P, dp, dp],
        "fcpt_totals_noalloc": [I, I, dp, dp, P, P, D, D, P, dp, dp],
        "fcpt_run_steps": [P, I, I],
        "fcpt_time_steps": [P, I, I, dp],
        "fcpt_time_kernel": [P, I, I, dp],
        "fcpt_time_copy": [P, I, dp, dp],
        "fcpt_set_launch_param": [P, I, I],
        "fcpt_get_launch_param": [P, I, C.POINTER(I)],
        "fcpt_rebase_backgrounds": [P],
        "fcpt_autotune": [P, I, C.POINTER(I), C.POINTER(I)],
        "fcpt_autotune_walks": [P, I, C.POINTER(I), C.POINTER(I)],
        "fcpt_disk_on_planet_accel": [P, C.POINTER(ForceBody), I, I, dp],
        "fcpt_batch_create": [C.POINTER(P), I, C.POINTER(P)],
        "fcpt_batch_destroy": [P],
        "fcpt_batch_synchronize": [P],
        "fcpt_batch_set_radial_waves": [P, I],
        "fcpt_batch_autotune": [P, I, C.POINTER(I), C.POINTER(I)],
        "fcpt_batch_run_steps": [P, I, I],
        "fcpt_batch_time_steps": [P, I, I, dp],
        "fcpt_batch_time_kernel": [P, I, I, dp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.fcpt_last_error.argtypes = [P]
    lib.fcpt_last_error.restype = C.c_char_p
    lib.fcpt_batch_last_error.argtypes = [P]
    lib.fcpt_batch_last_error.restype = C.c_char_p
    lib.fcpt_noalloc_last_error.argtypes = []
    lib.fcpt_noalloc_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def totals_noalloc(nr, nphi, rmed, surf, sigma_dev, vazi_dev, omega_frame, radius, stream=None):
    """fcpt_totals_noalloc over caller-owned FP64 device arrays (device
    addresses as ints); returns (mass, angular momentum)."""
    lib = load_library()
    rmed = np.ascontiguousarray(rmed, dtype=np.float64)
    surf = np.ascontiguousarray(surf, dtype=np.float64)
    m, l = C.c_double(), C.c_double()
    rc = lib.fcpt_totals_noalloc(nr, nphi, _dptr(rmed), _dptr(surf), sigma_dev, vazi_dev,
                                 omega_frame, radius, stream, C.byref(m), C.byref(l))
    if rc != 0:
        raise RuntimeError(f"fcpt_totals_noalloc -> {rc}: {lib.fcpt_noalloc_last_error().decode()}")
    return m.value, l.value


def exported_symbols() -> list[str]:
    """Function names declared in include/fcpt.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "fcpt.h")
    return re.findall(r"^\w[\w\s\*]*?\b(fcpt_\w+)\s*\(", open(hdr).read(), re.M)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Disk:
    """One disk resident on the GPU, driven through the C ABI."""

    def __init__(self, cfg: dict):
        self.lib = load_library()
        self.desc, self._radii = desc_from_dict(cfg)
        self.nr, self.nphi = self.desc.nr, self.desc.nphi
        h = C.c_void_p()
        self._check(self.lib.fcpt_create(C.byref(self.desc), C.byref(h)), None)
        self.h = h

    def _check(self, rc, h=0):
        if rc != 0:
            msg = self.lib.fcpt_last_error(self.h if h == 0 else h)
            raise RuntimeError(f"fcpt call failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fcpt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state
    def upload_state(self, sigma, vrad, vazi, energy=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (sigma, vrad, vazi)]
        e = None if energy is None else np.ascontiguousarray(energy, dtype=np.float64)
        self._check(self.lib.fcpt_upload_state(self.h, *[_dptr(a) for a in arrs],
                                               _dptr(e) if e is not None else None))

    def upload_reference(self, sigma0, vrad0, vazi0, energy0):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (sigma0, vrad0, vazi0, energy0)]
        self._check(self.lib.fcpt_upload_reference(self.h, *[_dptr(a) for a in arrs]))

    def download_state(self):
        """(sigma, vrad, vazi, energy) as FP64 host arrays (energy None when
        isothermal)."""
        out = [np.empty(field_shape(f, self.nr, self.nphi)) for f in ("SIGMA", "VRAD", "VAZI")]
        e = np.empty(field_shape("ENERGY", self.nr, self.nphi)) if self.desc.adiabatic else None
        self._check(self.lib.fcpt_download_state(self.h, *[_dptr(a) for a in out],
                                                 _dptr(e) if e is not None else None))
        return (*out, e)

    def upload_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == field_shape(name, self.nr, self.nphi), (name, a.shape)
        self._check(self.lib.fcpt_upload_field(self.h, FIELD_ID[name], _dptr(a)))

    def download_field(self, name):
        a = np.empty(field_shape(name, self.nr, self.nphi))
        self._check(self.lib.fcpt_download_field(self.h, FIELD_ID[name], _dptr(a)))
        return a

    def set_time(self, time, last_dt, n_monitor=0):
        self._check(self.lib.fcpt_set_time(self.h, time, last_dt, n_monitor))

    def get_time(self):
        t, l, n = C.c_double(), C.c_double(), C.c_int32()
        self._check(self.lib.fcpt_get_time(self.h, C.byref(t), C.byref(l), C.byref(n)))
        return t.value, l.value, n.value

    def set_step_params(self, planets, indirect=(0.0, 0.0)):
        sp = make_step_params(planets, indirect)
        self._check(self.lib.fcpt_set_step_params(self.h, C.byref(sp)))

    def init_profiles(self, **ic):
        """The reference's profile initial condition (init_physics), built on
        the device; keywords as make_ic."""
        d = make_ic(**ic)
        self._check(self.lib.fcpt_init_profiles(self.h, C.byref(d)))

    def write_snapshot(self, directory, snapshot=0, omega_frame=0.0, frame_angle=0.0, n_iter=0,
                       fields=()):
        """The reference's snapshot files in an existing directory; `fields`
        lists field names (empty: the restart set)."""
        info = SnapshotInfo(snapshot, 0, 0.0, omega_frame, frame_angle, 0.0, n_iter)
        mask = 0
        for f in fields:
            mask |= 1 << FIELD_ID[f]
        self._check(self.lib.fcpt_write_snapshot(self.h, os.fsencode(directory), C.byref(info), mask))

    def read_snapshot(self, directory):
        """Restart from a snapshot directory; returns its misc.bin record."""
        info = SnapshotInfo()
        self._check(self.lib.fcpt_read_snapshot(self.h, os.fsencode(directory), C.byref(info)))
        return {k: getattr(info, k) for k, _ in SnapshotInfo._fields_}

    def read_snapshot_reference(self, directory):
        self._check(self.lib.fcpt_read_snapshot_reference(self.h, os.fsencode(directory)))

    def write_grid_files(self, outdir, radial_spacing="Logarithmic"):
        self._check(self.lib.fcpt_write_grid_files(self.h, os.fsencode(outdir),
                                                   radial_spacing.encode()))

    def set_midstep_params(self, planets, indirect=(0.0, 0.0)):
        """Bodies and indirect term of a LeapFrog step's second kick."""
        sp = make_step_params(planets, indirect)
        self._check(self.lib.fcpt_set_midstep_params(self.h, C.byref(sp)))

    def compute_pressure(self):
        self._check(self.lib.fcpt_compute_pressure(self.h))

    # hydro step functions (names follow the reference)
    def condition_cfl(self):
        v = C.c_double()
        self._check(self.lib.fcpt_condition_cfl(self.h, C.byref(v)))
        return v.value

    def calculate_timestep(self):
        v = C.c_double()
        self._check(self.lib.fcpt_calculate_timestep(self.h, C.byref(v)))
        return v.value

    def nbody_potential(self):
        self._check(self.lib.fcpt_nbody_potential(self.h))

    def update_with_sourceterms(self, dt):
        self._check(self.lib.fcpt_update_with_sourceterms(self.h, dt))

    def artificial_viscosity(self, dt):
        self._check(self.lib.fcpt_artificial_viscosity(self.h, dt))

    def recalculate_viscosity(self):
        self._check(self.lib.fcpt_recalculate_viscosity(self.h))

    def viscous_stress_tensor(self):
        self._check(self.lib.fcpt_viscous_stress_tensor(self.h))

    def update_velocities_with_viscosity(self, dt):
        self._check(self.lib.fcpt_update_velocities_with_viscosity(self.h, dt))

    def substep3(self, time, dt):
        self._check(self.lib.fcpt_substep3(self.h, time, dt))

    def apply_boundary_condition(self, dt, final):
        self._check(self.lib.fcpt_apply_boundary_condition(self.h, dt, int(final)))

    def transport(self, dt):
        self._check(self.lib.fcpt_transport(self.h, dt))

    def recalculate_derived(self):
        self._check(self.lib.fcpt_recalculate_derived(self.h))

    def nshift(self):
        a = np.empty(self.nr, dtype=np.int32)
        self._check(self.lib.fcpt_get_nshift(self.h, a.ctypes.data_as(C.POINTER(C.c_int32))))
        return a

    def step_euler(self, dt):
        self._check(self.lib.fcpt_step_euler(self.h, dt))

    def step_leapfrog(self, dt):
        self._check(self.lib.fcpt_step_leapfrog(self.h, dt))

    def run_step(self):
        v = C.c_double()
        self._check(self.lib.fcpt_run_step(self.h, C.byref(v)))
        return v.value

    def totals(self, radius):
        m, l = C.c_double(), C.c_double()
        self._check(self.lib.fcpt_totals(self.h, radius, C.byref(m), C.byref(l)))
        return m.value, l.value

    def ring_geometry(self):
        """Rmed and Surf of rings 0 .. nr-1 as the context computed them."""
        rmed, surf = np.empty(self.nr), np.empty(self.nr)
        self._check(self.lib.fcpt_get_ring_geometry(self.h, _dptr(rmed), _dptr(surf)))
        return rmed, surf

    def synchronize(self):
        self._check(self.lib.fcpt_synchronize(self.h))

    def run_steps(self, n, use_graph=True):
        self._check(self.lib.fcpt_run_steps(self.h, n, int(use_graph)))

    def time_steps(self, n, use_graph=True):
        ms = C.c_double()
        self._check(self.lib.fcpt_time_steps(self.h, n, int(use_graph), C.byref(ms)))
        return ms.value

    def time_copy(self, reps):
        ms, nb = C.c_double(), C.c_double()
        self._check(self.lib.fcpt_time_copy(self.h, reps, C.byref(ms), C.byref(nb)))
        return ms.value, nb.value

    def time_kernel(self, kernel, reps):
        ms = C.c_double()
        self._check(self.lib.fcpt_time_kernel(self.h, kernel, reps, C.byref(ms)))
        return ms.value

    def disk_on_planet_accel(self, bodies, flags=FORCE_CORRECT_SELFGRAVITY):
        """ComputeDiskOnPlanetAccel for every body (x, y, r, r_cubic);
        returns an (n, 2) array of accelerations."""
        arr = make_force_bodies(bodies)
        out = np.zeros(2 * max(1, len(bodies)))
        self._check(self.lib.fcpt_disk_on_planet_accel(self.h, arr, len(bodies), flags, _dptr(out)))
        return out[:2 * len(bodies)].reshape(-1, 2)

    def set_fused_step(self, fused):
        """Fused source / artificial-viscosity / viscous passes in the step
        (default) or the separate kernels of the per-function entry points."""
        self._check(self.lib.fcpt_set_launch_param(self.h, PARAM_FUSED_STEP, 1 if fused else 0))

    def set_radial_tile(self, rings):
        """Rings walked by one radial-sweep thread (0: size-based default)."""
        self._check(self.lib.fcpt_set_launch_param(self.h, PARAM_RADIAL_TILE, rings))

    def set_visc_tile(self, rings):
        """Rings walked by one thread of the fused viscous pass (0:
        size-based default)."""
        self._check(self.lib.fcpt_set_launch_param(self.h, PARAM_VISC_TILE, rings))

    def set_kick_tile(self, rings):
        """Rings walked by one thread of the kick walk (0: size-based
        default; 1: the sources and artificial viscosity as separate
        passes)."""
        self._check(self.lib.fcpt_set_launch_param(self.h, PARAM_KICK_TILE, rings))

    def set_split_rounding(self, mode):
        """Split storage: 1 (default) rounds the evolved fields' deviations
        stochastically and the derived fields' to nearest, 2 every field
        stochastically, 0 every field to nearest."""
        self._check(self.lib.fcpt_set_launch_param(self.h, PARAM_SPLIT_ROUNDING, int(mode)))

    def set_rebase_interval(self, steps):
        """Split storage: move the ring backgrounds to the current ring means
        after every `steps` steps of run_steps (default 1000; 0: only on
        request)."""
        self._check(self.lib.fcpt_set_launch_param(self.h, PARAM_REBASE_INTERVAL, int(steps)))

    def rebase_backgrounds(self):
        """Split storage: move the ring backgrounds to the current ring means."""
        self._check(self.lib.fcpt_rebase_backgrounds(self.h))

    def launch_param(self, param):
        """The value launch parameter `param` (PARAM_*) takes in the next step."""
        v = C.c_int32()
        self._check(self.lib.fcpt_get_launch_param(self.h, param, C.byref(v)))
        return v.value

    def autotune(self, reps=10):
        """Time the tunable launch shapes on the current state and keep the
        fastest; returns {"radial_tile": t, "radial_waves": w}."""
        t, w = C.c_int32(), C.c_int32()
        self._check(self.lib.fcpt_autotune(self.h, reps, C.byref(t), C.byref(w)))
        return {"radial_tile": t.value, "radial_waves": w.value}

    def autotune_walks(self, reps=5):
        """Time the kick and viscous walks' tiles over whole steps (the state
        is put back after each trial) and keep the fastest; returns
        {"kick_tile": k, "visc_tile": v}."""
        k, v = C.c_int32(), C.c_int32()
        self._check(self.lib.fcpt_autotune_walks(self.h, reps, C.byref(k), C.byref(v)))
        return {"kick_tile": k.value, "visc_tile": v.value}

    def set_radial_waves(self, waves):
        """Register budget of the radial sweep in wavefronts per SIMD (2 or 3)."""
        self._check(self.lib.fcpt_set_launch_param(self.h, PARAM_RADIAL_WAVES, waves))


class Batch:
    """Many Disk objects advanced together, one launch per stage of the step
    (fcpt_batch_* in include/fcpt.h).  The disks keep their own state; read
    them with their download/totals methods between batch calls."""

    def __init__(self, disks):
        self.lib = load_library()
        self.disks = list(disks)
        arr = (C.c_void_p * len(self.disks))(*[d.h.value for d in self.disks])
        h = C.c_void_p()
        rc = self.lib.fcpt_batch_create(arr, len(self.disks), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"fcpt_batch_create failed ({rc})")
        self.h = h

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.fcpt_batch_last_error(self.h)
            raise RuntimeError(f"fcpt batch call failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fcpt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, reps=10):
        """Time the batched radial sweep's tiles and register budgets on the
        current state and keep the fastest; returns {"radial_tile": t,
        "radial_waves": w}."""
        t, w = C.c_int32(), C.c_int32()
        self._check(self.lib.fcpt_batch_autotune(self.h, reps, C.byref(t), C.byref(w)))
        return {"radial_tile": t.value, "radial_waves": w.value}

    def set_radial_waves(self, waves):
        self._check(self.lib.fcpt_batch_set_radial_waves(self.h, waves))

    def synchronize(self):
        self._check(self.lib.fcpt_batch_synchronize(self.h))

    def run_steps(self, n, use_graph=True):
        self._check(self.lib.fcpt_batch_run_steps(self.h, n, int(use_graph)))

    def time_steps(self, n, use_graph=True):
        ms = C.c_double()
        self._check(self.lib.fcpt_batch_time_steps(self.h, n, int(use_graph), C.byref(ms)))
        return ms.value

    def time_kernel(self, kernel, reps):
        ms = C.c_double()
        self._check(self.lib.fcpt_batch_time_kernel(self.h, kernel, reps, C.byref(ms)))
        return ms.value
