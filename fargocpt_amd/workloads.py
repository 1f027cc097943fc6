"""Disk configurations of the hydro-step workloads and a driver-side setup.

Code units are the reference's default (l0 = 1 au, m0 = 1 solar mass, so
G = 1; units.cpp:170-173).  The constants below are the code-unit values the
reference computes for that system with its LLNL units library
(constants.cpp:332-338), as recorded from the reference build: the gas
constant R, the Stefan-Boltzmann constant, the speed of light, Sigma0 = 200
g/cm^2 and the 3 K / 1e100 K temperature limits.
"""
from __future__ import annotations

from .ic import DiskSetup, initial_state, log_radii

CODE_UNITS = {
    "G": 1.0,
    "R": 1.0000000000000004,
    "sigma_sb": 468.30692052083845,
    "c_light": 10065.167938321798,
    "sigma0": 2.250928896887951e-05,   # 200 g/cm^2
    "t_min": 2.811616511825056e-05,    # 3 K
    "t_max": 9.372055039416855e+94,    # 1e100 K
}

# name -> (nr, nphi, rmax, adiabatic, artificial viscosity, beta cooling, planets [(a, q)])
WORKLOADS = {
    "iso_128x384": (128, 384, 2.5, False, 1, False, [(1.0, 1e-3)]),
    "adi_512x1536": (512, 1536, 2.5, True, 2, False, [(1.0, 1e-3)]),
    "iso_2048x6144": (2048, 6144, 2.5, False, 1, False, [(1.0, 1e-3)]),
    "adi_beta_4096x12288": (4096, 12288, 4.0, True, 2, True,
                            [(1.0, 1e-3), (1.6, 3e-4), (2.5, 1e-4)]),
}
DESCRIPTION = {
    "iso_128x384": "locally isothermal disk, 128x384 log rings in [0.4, 2.5], h 0.05, alpha 1e-3, "
                   "TW artificial viscosity, planet q 1e-3 at r 1, FARGO, CFL 0.5",
    "adi_512x1536": "adiabatic disk (gamma 1.4, energy equation, viscous heating), 512x1536, "
                    "alpha 1e-3, von Neumann-Richtmyer artificial viscosity, planet q 1e-3 at r 1",
    "iso_2048x6144": "locally isothermal disk of iso_128x384 at 2048x6144",
    "adi_beta_4096x12288": "adiabatic disk with beta cooling (beta 10), 4096x12288 in [0.4, 4.0], "
                           "planets q 1e-3, 3e-4, 1e-4 at r 1.0, 1.6, 2.5",
}


def disk_config(name: str, precision: int = 0) -> dict:
    """Descriptor fields of a workload; precision 0 = FP64 storage, 1 = split."""
    nr, nphi, rmax, adi, artvisc, beta, _ = WORKLOADS[name]
    return {
        "nr": nr, "nphi": nphi, "radii": log_radii(nr, 0.4, rmax), "rmin": 0.4, "rmax": rmax,
        "adiabatic": int(adi), "precision": int(precision),
        "G": CODE_UNITS["G"], "R": CODE_UNITS["R"], "sigma_sb": CODE_UNITS["sigma_sb"],
        "c_light": CODE_UNITS["c_light"], "hydro_center_mass": 1.0, "mu": 2.35, "gamma": 1.4,
        "aspect_ratio": 0.05, "flaring_index": 0.0, "alpha": 1e-3, "thickness_smoothing": 0.6,
        "cfl": 0.5, "cfl_max_var": 1.1, "heating_cooling_cfl_limit": 1.0,
        "sigma0": CODE_UNITS["sigma0"], "sigma_floor": 1e-9,
        "t_min": CODE_UNITS["t_min"], "t_max": CODE_UNITS["t_max"],
        "artificial_viscosity": artvisc, "artvisc_dissipation": 1, "artvisc_factor": 1.41,
        "heating_viscous": 1, "heating_viscous_factor": 1.0,
        "cooling_beta_enabled": int(beta), "cooling_beta_ref": 3, "cooling_beta": 10.0,
        "cooling_beta_ramp_up": 0.0, "radial_viscosity_factor": 1.0,
        "bc_inner": 0, "bc_outer": 0, "damping": 1,
        "damping_inner_limit": 1.1, "damping_outer_limit": 0.9, "damping_time_factor": 0.1,
        "damping_time_radius_outer": rmax,
        "damp_inner_scalar": -1, "damp_inner_vector": -1, "damp_outer_scalar": -1,
        "damp_outer_vector": -1, "omega_frame": 1.0, "monitor_timestep": 0.628,
    }


def setup(name: str, seed: int = 1337) -> DiskSetup:
    nr, nphi, rmax, adi, _, _, _ = WORKLOADS[name]
    return DiskSetup(nr=nr, nphi=nphi, rmin=0.4, rmax=rmax, sigma0=CODE_UNITS["sigma0"],
                     sigma_slope=0.5, aspect_ratio=0.05, flaring=0.0, G=1.0, mstar=1.0,
                     omega_frame=1.0, adiabatic=adi, gamma=1.4, noise_amp=1e-4, seed=seed)


def planet_inputs(name: str):
    """Planets on fixed circular positions (phi = 0) and the indirect term of
    the star-centred frame, -sum_k G m_k r_k / |r_k|^3 (the acceleration of the
    central star by the planets, Pframeforce.cpp:80-83 convention)."""
    _, _, _, _, _, _, planets = WORKLOADS[name]
    bodies = [(0.0, 0.0, 1.0, 0.0)]
    ix = iy = 0.0
    for a, q in planets:
        bodies.append((a, 0.0, q, 0.0))
        ix -= CODE_UNITS["G"] * q * a / (a ** 3)
    return bodies, (ix, iy)


def start(disk, name: str, first_dt: float = 0.1):
    """Upload the seeded initial condition and replay the reference's start-up
    calls (main.cpp:117, then sim::init, simulation.cpp:462-482)."""
    ic = initial_state(setup(name))
    disk.upload_reference(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    disk.upload_state(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    disk.set_time(0.0, first_dt, 0)
    disk.calculate_timestep()
    disk.apply_boundary_condition(0.0, False)
    disk.calculate_timestep()
    bodies, ind = planet_inputs(name)
    disk.set_step_params(bodies, ind)
    return ic


# The batched sweep: n locally isothermal disks (physics of iso_128x384) with
# Nr uniform in [128, 1024], Nphi = 3 Nr, planet mass ratio log-uniform in
# [1e-5, 1e-3], alpha log-uniform in [1e-4, 1e-2], h uniform in [0.03, 0.07],
# drawn from numpy's PCG64 with the given seed; disk k's noise seed is seed + k.
BATCHES = {"batch64": (64, 1337)}


def batch_members(name: str = "batch64") -> list:
    import numpy as np
    n, seed = BATCHES[name]
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nr = int(rng.integers(128, 1025))
        q = float(10.0 ** rng.uniform(-5.0, -3.0))
        alpha = float(10.0 ** rng.uniform(-4.0, -2.0))
        h = float(rng.uniform(0.03, 0.07))
        out.append({"nr": nr, "nphi": 3 * nr, "q": q, "alpha": alpha, "h": h, "seed": seed + k})
    return out


def member_config(m: dict, precision: int = 0) -> dict:
    """Descriptor of one batched disk: iso_128x384's physics at the member's
    size, viscosity and aspect ratio."""
    cfg = disk_config("iso_128x384", precision)
    cfg.update({"nr": m["nr"], "nphi": m["nphi"], "radii": log_radii(m["nr"], 0.4, 2.5),
                "alpha": m["alpha"], "aspect_ratio": m["h"]})
    return cfg


def member_start(disk, m: dict, first_dt: float = 0.1):
    """Initial state and start-up calls of one batched disk (as start())."""
    setup_m = DiskSetup(nr=m["nr"], nphi=m["nphi"], rmin=0.4, rmax=2.5, sigma0=CODE_UNITS["sigma0"],
                        sigma_slope=0.5, aspect_ratio=m["h"], flaring=0.0, G=1.0, mstar=1.0,
                        omega_frame=1.0, adiabatic=False, gamma=1.4, noise_amp=1e-4,
                        seed=m["seed"])
    ic = initial_state(setup_m)
    disk.upload_reference(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    disk.upload_state(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    disk.set_time(0.0, first_dt, 0)
    disk.calculate_timestep()
    disk.apply_boundary_condition(0.0, False)
    disk.calculate_timestep()
    q = m["q"]
    disk.set_step_params([(0.0, 0.0, 1.0, 0.0), (1.0, 0.0, q, 0.0)], (-CODE_UNITS["G"] * q, 0.0))
    return ic


def algorithmic_bytes(kernel: str, adiabatic: bool, element: int = 8) -> int:
    """Compulsory HBM bytes per cell of one launch of a step kernel: each
    stored field it reads once plus each field it writes once, `element` bytes
    per value (8 for FP64 storage, 4 for split storage)."""
    e = 1 if adiabatic else 0
    reads_writes = {
        "radial_sweep": (3 + e, 5 + e),        # Sigma, v_r, v_phi, (E) -> 4 momenta, Sigma, (E)
        "azimuthal_sweep": (6 + e, 5 + e),     # Sigma, 4 momenta, (E), v_phi -> 5/6, shifted
        "cfl": (4 + 3 * e, 0),                 # v_r, v_phi, c_s, nu, (Q+, Q-, E)
        "stress": (4, 4),                      # v_r, v_phi, nu, Sigma -> div v, tau_rr, tau_pp, tau_rp
        # v_r, v_phi, Sigma, (nu) -> v_r, v_phi, (div v and the tensor for the heating)
        "visc_fused": (3 + e, 2 + 4 * e),
        "potential": (1, 1),                   # H -> Phi
        "disk_force": (2 + e, 0),              # Sigma (ring means, then cells), (H)
    }
    r, w = reads_writes[kernel]
    return element * (r + w)


def step_bytes(adiabatic: bool, element: int = 8) -> int:
    """Evolved fields (Sigma, v_r, v_phi and E) read and written once per step."""
    return element * 2 * (4 if adiabatic else 3)


STEP_BYTES = {False: step_bytes(False), True: step_bytes(True)}


def cells(name: str) -> int:
    nr, nphi = WORKLOADS[name][:2]
    return nr * nphi


def is_adiabatic(name: str) -> bool:
    return WORKLOADS[name][3]


__all__ = ["CODE_UNITS", "WORKLOADS", "DESCRIPTION", "disk_config", "setup", "planet_inputs",
           "start", "algorithmic_bytes", "step_bytes", "STEP_BYTES", "cells", "is_adiabatic",
           "BATCHES", "batch_members", "member_config", "member_start"]
