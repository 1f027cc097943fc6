"""Helpers that turn the committed reference fixtures into driver calls.

Each tests/golden/<case>.npz was written by tests/golden/make_golden.py from
the reference FargoCPT build.  It holds the reference's code-unit constants and
parameters (meta), the geometry it computed, the per-step host inputs it fed
its potential (planet x, y, mass, cubic radius and the indirect term), the
per-step dt values and FARGO shifts, and checkpointed states.
"""
from __future__ import annotations

import json
import os

import numpy as np

from fargocpt_amd.ic import DiskSetup, initial_state

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# stage order of the reference's step_Euler as dumped by refdump.cpp
STAGES_ISO = ["start", "potential", "sources", "artvisc", "viscosity", "stress",
              "viscupdate", "bc", "transport", "bcfinal", "derived"]
STAGES_ADI = ["start", "potential", "sources", "artvisc", "viscosity", "stress",
              "viscupdate", "substep3", "bc", "transport", "bcfinal", "derived"]
STAGE_FIELDS = {
    "sigma": "SIGMA", "vrad": "VRAD", "vazi": "VAZI", "energy": "ENERGY",
    "pressure": "PRESSURE", "soundspeed": "SOUNDSPEED", "scaleheight": "SCALE_HEIGHT",
    "viscosity": "VISCOSITY", "potential": "POTENTIAL", "qr": "Q_R", "qphi": "Q_PHI",
    "divv": "DIV_V", "taurr": "TAU_R_R", "taurp": "TAU_R_PHI", "taupp": "TAU_PHI_PHI",
    "qplus": "QPLUS", "qminus": "QMINUS", "temperature": "TEMPERATURE",
}
FIRST_DT = 0.1  # FirstDT of every fixture config
BC_IDS = {"reflecting": 0, "zerogradient": 1, "outflow": 2}


def load_case(name: str):
    data = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    meta = json.loads(bytes(data["meta_json"]).decode())
    return meta, data


def disk_config(meta: dict) -> dict:
    """Flat parameter dict for fargocpt_amd.capi.desc_from_dict."""
    from fargocpt_amd.ic import log_radii
    nr = meta["Nr"]
    radii = log_radii(nr, meta["RMIN"], meta["RMAX"])
    return {
        "nr": nr, "nphi": meta["Nphi"], "radii": radii, "rmin": meta["RMIN"], "rmax": meta["RMAX"],
        "adiabatic": meta["adiabatic"], "precision": 0,
        "G": meta["G"], "R": meta["R"], "sigma_sb": meta["sigma_sb"], "c_light": meta["c"],
        "hydro_center_mass": meta["hydro_center_mass"], "mu": meta["mu"], "gamma": meta["gamma"],
        "aspect_ratio": meta["aspectratio"], "flaring_index": meta["flaring"],
        "alpha": meta["alpha"], "thickness_smoothing": meta["thickness_smoothing"],
        "cfl": meta["cfl"], "cfl_max_var": meta["cfl_max_var"],
        "heating_cooling_cfl_limit": meta.get("heating_cooling_cfl_limit", 1.0),
        "sigma0": meta["sigma0"], "sigma_floor": meta["sigma_floor"],
        "t_min": meta["Tmin"], "t_max": meta["Tmax"],
        "artificial_viscosity": meta["artvisc"],
        "artvisc_dissipation": meta["artvisc_dissipation"],
        "artvisc_factor": meta["artvisc_factor"],
        "heating_viscous": meta["heating_viscous"],
        "heating_viscous_factor": meta["heating_viscous_factor"],
        "cooling_beta_enabled": meta["cooling_beta_enabled"],
        "cooling_beta_ref": (3 if meta["cooling_beta_floor"]
                             else 1 if meta.get("cooling_beta_reference") else 0),
        "cooling_beta": meta["cooling_beta"],
        "cooling_beta_ramp_up": meta.get("cooling_beta_ramp_up", 0.0),
        "radial_viscosity_factor": meta["radial_viscosity_factor"],
        "bc_inner": BC_IDS[meta.get("bc_inner", "reflecting")],
        "bc_outer": BC_IDS[meta.get("bc_outer", "reflecting")],
        "damping": meta.get("damping_enabled", 1),
        "damping_inner_limit": meta["damping_inner_limit"],
        "damping_outer_limit": meta["damping_outer_limit"],
        "damping_time_factor": meta["damping_time_factor"],
        "damping_time_radius_outer": meta["damping_time_radius_outer"],
        "damp_inner_scalar": -1, "damp_inner_vector": -1,
        "damp_outer_scalar": -1, "damp_outer_vector": -1,
        "omega_frame": meta["OmegaFrame"], "monitor_timestep": meta["monitor_timestep"],
        "integrator": meta.get("integrator", 0),
        "potential_smoothing": ((1 if meta.get("compatibility_no_star_smoothing", 0) else 0) |
                                (2 if meta.get("compatibility_smoothing_planetloc", 0) else 0)),
    }


def initial_condition(meta: dict) -> dict:
    setup = DiskSetup(nr=meta["Nr"], nphi=meta["Nphi"], rmin=meta["RMIN"], rmax=meta["RMAX"],
                      sigma0=meta["sigma0"], sigma_slope=meta["sigma_slope"],
                      aspect_ratio=meta["aspectratio"], flaring=meta["flaring"], G=meta["G"],
                      mstar=meta["hydro_center_mass"], omega_frame=meta["OmegaFrame"],
                      adiabatic=bool(meta["adiabatic"]), gamma=meta["gamma"],
                      noise_amp=meta["noise_amp"], seed=meta["seed"])
    return initial_state(setup)


def start_run(disk, meta: dict):
    """Upload the initial condition and replay the reference's start-up calls
    (main.cpp:117 CalculateTimeStep, then sim::init, simulation.cpp:462-482)."""
    ic = initial_condition(meta)
    disk.upload_reference(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    disk.upload_state(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    disk.set_time(0.0, FIRST_DT, 0)
    disk.calculate_timestep()
    disk.apply_boundary_condition(0.0, False)
    disk.calculate_timestep()
    return ic


def step_inputs(data, step: int):
    """(planets, indirect) the reference's potential saw at 1-based step."""
    pl = data["planets"][step - 1]
    planets = [tuple(p) for p in pl]
    rec = data["steps"][step - 1]
    return planets, (rec[6], rec[7])


IC_FIELDS = [("sigma", "SIGMA"), ("vrad", "VRAD"), ("vazi", "VAZI"), ("energy", "ENERGY"),
             ("qplus", "QPLUS"), ("qminus", "QMINUS"), ("pressure", "PRESSURE"),
             ("soundspeed", "SOUNDSPEED"), ("temperature", "TEMPERATURE"),
             ("viscosity", "VISCOSITY")]
IC_REFERENCE = [("sigma0", "SIGMA0"), ("vrad0", "VRAD0"), ("vazi0", "VAZI0"),
                ("energy0", "ENERGY0")]


def ic_params(meta: dict) -> dict:
    """Keywords of init_profiles for an initial-condition fixture."""
    return dict(meta["ic"], sigma_slope=meta["sigma_slope"])


def step_inputs_mid(data, step: int):
    """(planets, indirect) of a LeapFrog step's second kick, or None."""
    if "planets_mid" not in data.files:
        return None
    pl = data["planets_mid"][step - 1]
    rec = data["steps"][step - 1]
    return [tuple(p) for p in pl], (rec[10], rec[11])


def stage_states(data, meta: dict, step: int):
    """Yield (stage_name, full_state_dict) of the per-stage dumps of one step."""
    stages = STAGES_ADI if meta["adiabatic"] else STAGES_ISO
    cur = {}
    # the first stage of step 1 holds every field; later stages only changes
    for s in range(1, step + 1):
        for k, name in enumerate(stages):
            for f in STAGE_FIELDS:
                key = f"stage_{s}_{k:02d}_{name}_{f}"
                if key in data.files:
                    cur[f] = data[key]
            if s == step:
                yield name, dict(cur)


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a-b| / max |b| (0 when both are zero)."""
    scale = float(np.max(np.abs(b)))
    diff = float(np.max(np.abs(a - b)))
    if scale == 0.0:
        return diff
    return diff / scale


def call_stage(disk, stage: str, dt: float, t_before: float):
    """Invoke the driver function that produced one per-stage dump."""
    if stage == "potential":
        disk.nbody_potential()
    elif stage == "sources":
        disk.update_with_sourceterms(dt)
    elif stage == "artvisc":
        disk.artificial_viscosity(dt)
    elif stage == "viscosity":
        disk.recalculate_viscosity()
    elif stage == "stress":
        disk.viscous_stress_tensor()
    elif stage == "viscupdate":
        disk.update_velocities_with_viscosity(dt)
    elif stage == "substep3":
        disk.substep3(t_before, dt)
    elif stage == "bc":
        disk.apply_boundary_condition(0.0, False)
    elif stage == "transport":
        disk.transport(dt)
    elif stage == "bcfinal":
        disk.apply_boundary_condition(dt, True)
    elif stage == "derived":
        disk.recalculate_derived()
    else:
        raise ValueError(stage)


def check_stages(disk, meta, data, step: int):
    """Replay every stage of one reference step from the dumped inputs.

    Returns (worst relative error per stage and field, nshift_equal)."""
    disk.upload_reference(data["init_sigma0"], data["init_vrad0"], data["init_vazi0"],
                          data["init_energy0"])
    rec = data["steps"][step - 1]
    dt, t = rec[3], rec[0]
    planets, ind = step_inputs(data, step)
    disk.set_step_params(planets, ind)
    disk.set_time(t, rec[1], int(rec[4]))
    errors = {}
    nshift_ok = None
    prev = None
    for stage, state in stage_states(data, meta, step):
        if prev is not None:
            for f, name in STAGE_FIELDS.items():
                disk.upload_field(name, prev[f])
            call_stage(disk, stage, dt, t)
            if stage == "transport":
                nshift_ok = bool(np.array_equal(disk.nshift(), data["nshift"][step - 1]))
            for f, name in STAGE_FIELDS.items():
                errors[(stage, f)] = rel_err(disk.download_field(name), state[f])
        prev = state
    return errors, nshift_ok


STATE_FIELDS = [("sigma", "SIGMA"), ("vrad", "VRAD"), ("vazi", "VAZI"), ("energy", "ENERGY"),
                ("qplus", "QPLUS"), ("qminus", "QMINUS")]


def compare_state(disk, data, tag: str) -> dict:
    """Relative error of the driver's state against a checkpoint (whole arrays,
    or per-ring sums / extrema / sample rings for large grids)."""
    out = {}
    for f, name in STATE_FIELDS:
        a = disk.download_field(name)
        key = f"state_{tag}_{f}"
        if key in data.files:
            out[f] = rel_err(a, data[key])
        elif key + "_ringsum" in data.files:
            out[f + "_ringsum"] = rel_err(a.sum(axis=1), data[key + "_ringsum"])
            out[f + "_ringmin"] = rel_err(a.min(axis=1), data[key + "_ringmin"])
            out[f + "_ringmax"] = rel_err(a.max(axis=1), data[key + "_ringmax"])
            rows = data[key + "_rowidx"]
            out[f + "_rows"] = rel_err(a[rows], data[key + "_rows"])
    return out


def run_case(disk, meta, data, nsteps=None, totals=False):
    """Replay a whole reference run: start-up calls, then sim::run iterations
    with the reference's per-step planet inputs.  Returns
    (max relative dt error, number of steps whose FARGO shifts differ,
     {checkpoint: errors}); with `totals`, also the relative errors of the
    disk's mass and angular momentum after the last step against the
    reference's gas_total_mass / gas_angular_momentum (whole disk and up to
    the middle ring's radius) under the key "totals"."""
    start_run(disk, meta)
    nsteps = nsteps or meta["nsteps"]
    res = {"0": compare_state(disk, data, "0")}
    worst_dt = 0.0
    bad_shift = 0
    for s in range(1, nsteps + 1):
        planets, ind = step_inputs(data, s)
        disk.set_step_params(planets, ind)
        mid = step_inputs_mid(data, s)
        if mid is not None:
            disk.set_midstep_params(*mid)
        dt = disk.run_step()
        ref_dt = data["steps"][s - 1][3]
        worst_dt = max(worst_dt, abs(dt - ref_dt) / ref_dt)
        if not np.array_equal(disk.nshift(), data["nshift"][s - 1]):
            bad_shift += 1
        if f"state_{s}_sigma" in data.files or f"state_{s}_sigma_ringsum" in data.files:
            res[str(s)] = compare_state(disk, data, str(s))
    if totals:
        ref = data["totals"][nsteps - 1]
        m, l = disk.totals(1e300)
        mc, lc = disk.totals(ref[4])
        res["totals"] = {"mass": abs(m - ref[0]) / abs(ref[0]),
                         "angmom": abs(l - ref[1]) / abs(ref[1]),
                         "mass_cut": abs(mc - ref[2]) / abs(ref[2]),
                         "angmom_cut": abs(lc - ref[3]) / abs(ref[3])}
    return worst_dt, bad_shift, res


def totals_magnitudes(sigma, vazi, surf, rmed, omega_frame, radius, nr):
    """Sums of |term| of the mass and angular-momentum totals over the active
    rings with Rmed <= radius (quantities.cpp:51-78, 242-279) and the number
    of terms: the scale of the rounding bound between summation orders."""
    rows = [i for i in range(1, nr - 1) if rmed[i] <= radius]
    if not rows:
        return 0.0, 0.0, 0
    r = np.array(rows)
    s = sigma[r]
    sm = 0.5 * (s + np.roll(s, 1, axis=1))
    dm = surf[r, None] * s
    dl = surf[r, None] * sm * rmed[r, None] * (vazi[r] + omega_frame * rmed[r, None])
    return float(np.sum(np.abs(dm))), float(np.sum(np.abs(dl))), dm.size


def totals_bound(n, magnitude):
    """|compensated - left-to-right| for n terms: (n - 1) u + 2 u of sum |term|,
    u = 2^-53, with a margin for the bound's own rounding."""
    return (n + 4) * 2.0 ** -53 * magnitude * 1.01
