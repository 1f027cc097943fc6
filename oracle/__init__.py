"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this module, as the reference the GPU results are compared with or as the
timed CPU baseline.  It exposes the same
method names as fargocpt_amd.capi.Disk so tests can drive both the same way.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from fargocpt_amd.capi import (FIELD_ID, FORCE_CORRECT_SELFGRAVITY, DiskDesc, ForceBody, IcDesc,
                               desc_from_dict, field_shape, make_force_bodies, make_ic,
                               make_step_params, StepParams)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P, D, I = C.c_void_p, C.c_double, C.c_int
    dp = C.POINTER(C.c_double)
    spec = {
        "fo_create": ([C.POINTER(DiskDesc)], P),
        "fo_destroy": ([P], None),
        "fo_set_threads": ([I], None),
        "fo_threads": ([], I),
        "fo_field": ([P, I], dp),
        "fo_field_rows": ([P, I], I),
        "fo_set_state": ([P, dp, dp, dp, dp], None),
        "fo_set_reference": ([P, dp, dp, dp, dp], None),
        "fo_set_time": ([P, D, D, I], None),
        "fo_get_time": ([P, dp, dp, C.POINTER(I)], None),
        "fo_set_step_params": ([P, C.POINTER(StepParams)], None),
        "fo_set_midstep_params": ([P, C.POINTER(StepParams)], None),
        "fo_init_profiles": ([P, C.POINTER(IcDesc)], None),
        "fo_step_leapfrog": ([P, D], None),
        "fo_compute_pressure": ([P], None),
        "fo_geom": ([P, C.c_char_p], dp),
        "fo_damping_ids": ([P, C.POINTER(I)], I),
        "fo_condition_cfl": ([P], D),
        "fo_calculate_timestep": ([P], D),
        "fo_nbody_potential": ([P], None),
        "fo_update_with_sourceterms": ([P, D], None),
        "fo_artificial_viscosity": ([P, D], None),
        "fo_recalculate_viscosity": ([P], None),
        "fo_viscous_stress_tensor": ([P], None),
        "fo_update_velocities_with_viscosity": ([P, D], None),
        "fo_substep3": ([P, D, D], None),
        "fo_apply_boundary_condition": ([P, D, I], None),
        "fo_transport": ([P, D], None),
        "fo_recalculate_derived": ([P], None),
        "fo_nshift": ([P], C.POINTER(C.c_int32)),
        "fo_step_euler": ([P, D], None),
        "fo_run_step": ([P], D),
        "fo_totals": ([P, D, dp, dp], None),
        "fo_disk_on_planet_accel": ([P, C.POINTER(ForceBody), I, I, dp], None),
    }
    for name, (args, res) in spec.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleDisk:
    """CPU oracle disk with the method names of fargocpt_amd.capi.Disk."""

    def __init__(self, cfg: dict):
        self.lib = load()
        self.desc, self._radii = desc_from_dict(cfg)
        self.nr, self.nphi = self.desc.nr, self.desc.nphi
        self.h = self.lib.fo_create(C.byref(self.desc))

    def close(self):
        if getattr(self, "h", None):
            self.lib.fo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _view(self, name):
        rows, cols = field_shape(name, self.nr, self.nphi)
        ptr = self.lib.fo_field(self.h, FIELD_ID[name])
        return np.ctypeslib.as_array(ptr, shape=(rows * cols,)).reshape(rows, cols)

    def upload_state(self, sigma, vrad, vazi, energy=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (sigma, vrad, vazi)]
        e = None if energy is None else np.ascontiguousarray(energy, dtype=np.float64)
        self.lib.fo_set_state(self.h, *[_dptr(a) for a in arrs], _dptr(e) if e is not None else None)

    def upload_reference(self, sigma0, vrad0, vazi0, energy0):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (sigma0, vrad0, vazi0, energy0)]
        self.lib.fo_set_reference(self.h, *[_dptr(a) for a in arrs])

    def upload_field(self, name, arr):
        self._view(name)[...] = arr

    def download_field(self, name):
        return self._view(name).copy()

    def geom(self, name, n):
        p = self.lib.fo_geom(self.h, name.encode())
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def damping_ids(self):
        a = (C.c_int * 4)()
        self.lib.fo_damping_ids(self.h, a)
        return list(a)

    def set_time(self, time, last_dt, n_monitor=0):
        self.lib.fo_set_time(self.h, time, last_dt, n_monitor)

    def get_time(self):
        t, l, n = C.c_double(), C.c_double(), C.c_int()
        self.lib.fo_get_time(self.h, C.byref(t), C.byref(l), C.byref(n))
        return t.value, l.value, n.value

    def set_step_params(self, planets, indirect=(0.0, 0.0)):
        sp = make_step_params(planets, indirect)
        self.lib.fo_set_step_params(self.h, C.byref(sp))

    def condition_cfl(self):
        return self.lib.fo_condition_cfl(self.h)

    def calculate_timestep(self):
        return self.lib.fo_calculate_timestep(self.h)

    def nbody_potential(self):
        self.lib.fo_nbody_potential(self.h)

    def update_with_sourceterms(self, dt):
        self.lib.fo_update_with_sourceterms(self.h, dt)

    def artificial_viscosity(self, dt):
        self.lib.fo_artificial_viscosity(self.h, dt)

    def recalculate_viscosity(self):
        self.lib.fo_recalculate_viscosity(self.h)

    def viscous_stress_tensor(self):
        self.lib.fo_viscous_stress_tensor(self.h)

    def update_velocities_with_viscosity(self, dt):
        self.lib.fo_update_velocities_with_viscosity(self.h, dt)

    def substep3(self, time, dt):
        self.lib.fo_substep3(self.h, time, dt)

    def apply_boundary_condition(self, dt, final):
        self.lib.fo_apply_boundary_condition(self.h, dt, int(final))

    def transport(self, dt):
        self.lib.fo_transport(self.h, dt)

    def recalculate_derived(self):
        self.lib.fo_recalculate_derived(self.h)

    def nshift(self):
        p = self.lib.fo_nshift(self.h)
        return np.ctypeslib.as_array(p, shape=(self.nr,)).copy()

    def step_euler(self, dt):
        self.lib.fo_step_euler(self.h, dt)

    def step_leapfrog(self, dt):
        self.lib.fo_step_leapfrog(self.h, dt)

    def compute_pressure(self):
        self.lib.fo_compute_pressure(self.h)

    def set_midstep_params(self, planets, indirect=(0.0, 0.0)):
        sp = make_step_params(planets, indirect)
        self.lib.fo_set_midstep_params(self.h, C.byref(sp))

    def init_profiles(self, **ic):
        d = make_ic(**ic)
        self.lib.fo_init_profiles(self.h, C.byref(d))

    def run_step(self):
        return self.lib.fo_run_step(self.h)

    def totals(self, radius):
        m, l = C.c_double(), C.c_double()
        self.lib.fo_totals(self.h, radius, C.byref(m), C.byref(l))
        return m.value, l.value

    def disk_on_planet_accel(self, bodies, flags=FORCE_CORRECT_SELFGRAVITY):
        arr = make_force_bodies(bodies)
        out = np.zeros(2 * max(1, len(bodies)))
        self.lib.fo_disk_on_planet_accel(self.h, arr, len(bodies), flags, _dptr(out))
        return out[:2 * len(bodies)].reshape(-1, 2)

    def synchronize(self):
        pass

    def set_threads(self, n: int):
        self.lib.fo_set_threads(n)

    def threads(self) -> int:
        return self.lib.fo_threads()
