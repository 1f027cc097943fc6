#!/usr/bin/env python3
"""Generate the committed golden fixtures from the reference FargoCPT build.

Runs only in a container that holds the reference sources at
/root/reference/src.  Steps:

1. build the reference objects (``make -f oracle/ref.mk``) and the fixture
   driver ``tests/golden/refdump.cpp`` linked against them (non-PIE, so the
   address of the reference's file-static FARGO shift table ``Nshift``,
   src/TransportEuler.cpp:49, can be read from the symbol table with ``nm``);
2. for every case below: write the YAML config, query the reference's code
   units (sigma0, G), write the package's seeded initial condition
   (fargocpt_amd/ic.py), run the driver, and pack its outputs into
   ``tests/golden/<case>.npz`` (float64 arrays, int32 shifts, meta as JSON);
3. run the Euler case c1_small and the LeapFrog cases a second time through
   the reference's own ``sim::run`` and check that the driver's step loops
   reproduce them bit for bit.

Usage:  python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_OUT = os.path.join(ROOT, "oracle", "_ref")
sys.path.insert(0, ROOT)

from fargocpt_amd.ic import DiskSetup, initial_state  # noqa: E402

BASE = {
    "Disk": "yes", "DiskFeedback": "no", "SelfGravity": "no", "IntegrateParticles": "no",
    "l0": "1 au", "m0": "1 solMass", "mu": "2.35",
    "HydroFrameCenter": "primary", "IndirectTermMode": "0", "OmegaFrame": "1.0", "Frame": "F",
    "MonitorTimestep": "0.628", "Nmonitor": "10", "Nsnapshots": "50", "FirstDT": "1.0e-1",
    "Transport": "FARGO", "Integrator": "Euler", "CFL": "0.5", "CFLmaxVar": "1.1", "cps": "-1.0",
    "Nrad": "128", "Naz": "384", "Rmin": "0.4", "Rmax": "2.5", "RadialSpacing": "Logarithmic",
    "ThicknessSmoothing": "0.6", "ThicknessSmoothingSG": "0.6",
    "Sigma0": "200 g/cm2", "SigmaSlope": "0.5", "SigmaFloor": "1e-9",
    "AspectRatio": "0.05", "FlaringIndex": "0.0", "AspectRatioMode": "0", "RandomSigma": "'No'",
    "ViscousAlpha": "1.0e-3", "ArtificialViscosity": "TW", "ArtificialViscosityDissipation": "Yes",
    "ArtificialViscosityFactor": "1.41", "EquationOfState": "isothermal", "AdiabaticIndex": "1.4",
    "SurfaceCooling": "no", "RadiativeDiffusion": "No", "CoolingBetaLocal": "no",
    "CoolingBeta": "10", "CoolingBetaReference": "floor",
    "HeatingViscous": "Yes", "MinimumTemperature": "3 K", "MaximumTemperature": "1e100 K",
    "HeatingCoolingCFLlimit": "1.0", "InnerBoundary": "Reflecting", "OuterBoundary": "Reflecting",
    "Damping": "Yes", "DampingInnerLimit": "1.10", "DampingOuterLimit": "0.90",
    "DampingTimeFactor": "1.0e-1", "DampingTimeRadiusOuter": "2.5",
    "DampingEnergyInner": "Initial", "DampingVRadialInner": "Initial",
    "DampingVAzimuthalInner": "Initial", "DampingSurfaceDensityInner": "Initial",
    "DampingEnergyOuter": "Initial", "DampingVRadialOuter": "Initial",
    "DampingVAzimuthalOuter": "Initial", "DampingSurfaceDensityOuter": "Initial",
    "OutputDir": "output/run", "LogAfterRealSeconds": "600", "LogAfterSteps": "0",
    "WriteAtEveryTimestep": "No", "WriteDensity": "Yes", "WriteVelocity": "Yes",
    "WriteEnergy": "Yes", "WriteDiskQuantities": "No",
}

PLANET_1 = [(1.0, 1e-3)]
PLANETS_3 = [(1.0, 1e-3), (1.6, 3e-4), (2.5, 1e-4)]
# (a, q, cubic smoothing factor, ramp-up time in orbits)
PLANETS_MIX = [(1.0, 1e-3, 1.5, 0.0), (1.5, 5e-4, 0.0, 0.5)]

ADIABATIC = {"EquationOfState": "Ideal", "ArtificialViscosity": "SN"}

# name: (overrides, planets, nsteps, stage_steps, checkpoints, full_final_state)
CASES = {
    # locally isothermal 128 x 384, 100 steps (the CPU-runnable configuration)
    "c1_iso": ({}, PLANET_1, 100, 0, [1, 10, 100], True),
    # same physics on a 32 x 64 mesh, per-stage dumps of the first two steps
    "c1_small": ({"Nrad": "32", "Naz": "64"}, PLANET_1, 20, 2, [1, 5, 20], True),
    # adiabatic (gamma 1.4, von Neumann-Richtmyer viscosity, viscous heating) 32 x 64
    "c2_small": (dict(ADIABATIC, Nrad="32", Naz="64"), PLANET_1, 20, 2, [1, 5, 20], True),
    # adiabatic with beta cooling (beta 10, floor reference) and three planets, 48 x 144
    "c5_small": (dict(ADIABATIC, Nrad="48", Naz="144", Rmax="4.0", DampingTimeRadiusOuter="4.0",
                      CoolingBetaLocal="yes"), PLANETS_3, 20, 1, [1, 5, 20], True),
    # branches the cases above leave out: adiabatic with TW viscosity, outflow inner and
    # zero-gradient outer boundaries, Klahr-Kley cubic smoothing, beta cooling towards
    # the reference state with a ramp-up, a planet mass ramp-up; 32 x 64
    "mix_small": (dict(EquationOfState="Ideal", ArtificialViscosity="TW", Nrad="32", Naz="64",
                       InnerBoundary="Outflow", OuterBoundary="ZeroGradient",
                       CoolingBetaLocal="yes", CoolingBetaReference="reference",
                       CoolingBetaRampUp="1.0"), PLANETS_MIX, 20, 1, [1, 5, 20], True),
    # locally isothermal without artificial viscosity and without damping, zero-gradient
    # boundaries; 32 x 64
    "none_small": (dict(ArtificialViscosity="none", Nrad="32", Naz="64", Damping="No",
                        InnerBoundary="ZeroGradient", OuterBoundary="ZeroGradient"),
                   PLANET_1, 20, 1, [1, 5, 20], True),
    # the potential smoothing switches of compute_smoothing (Force.cpp:145-159): both
    # (unsmoothed star, planet smoothing fixed at its location), isothermal 32 x 64; and
    # the unsmoothed star alone on an adiabatic 32 x 64 disk (the planet's smoothing
    # from the 2-D scale height)
    "smooth_small": ({"Nrad": "32", "Naz": "64", "CompatibilityNoStarSmoothing": "yes",
                      "CompatibilitySmoothingPlanetLoc": "yes"}, PLANET_1, 20, 1, [1, 5, 20], True),
    "nostar_small": (dict(ADIABATIC, Nrad="32", Naz="64", CompatibilityNoStarSmoothing="yes"),
                     PLANET_1, 20, 1, [1, 5, 20], True),
    # adiabatic 512 x 1536, 500 steps: only per-ring moments and a few full rings kept
    "c2_adi": (dict(ADIABATIC, Nrad="512", Naz="1536"), PLANET_1, 500, 0, [500], False),
    # the LeapFrog integrator (kick dt/2, drift dt, kick dt/2): the isothermal 32 x 64
    # disk, and an adiabatic one with beta cooling (ramp-up, so SubStep3's time
    # matters), a cubic-smoothed planet and a planet whose mass ramps up (the
    # two kicks see different bodies)
    "lf_iso": ({"Nrad": "32", "Naz": "64", "Integrator": "Leapfrog"}, PLANET_1, 20, 1,
               [1, 5, 20], True),
    "lf_adi": (dict(ADIABATIC, Nrad="32", Naz="64", Integrator="Leapfrog", CoolingBetaLocal="yes",
                    CoolingBetaRampUp="1.0"), PLANETS_MIX, 20, 1, [1, 5, 20], True),
}


# The reference's own initial-condition generators (init_physics): name ->
# (overrides, planets).  A 32 x 64 mesh; the fixture holds the state right
# after init_physics.
IC_CASES = {
    # power-law Sigma, pressure-supported smoothed v_phi, numerical viscous v_r
    "ic_iso": ({"Nrad": "32", "Naz": "64"}, PLANET_1),
    # adiabatic: energy profile with the temperature floor
    "ic_adi": (dict(ADIABATIC, Nrad="32", Naz="64", SigmaSlope="1.0", FlaringIndex="0.25"),
               PLANET_1),
    # exponential cutoffs at both edges down to a raised floor, an imposed disk
    # drift, flaring
    "ic_cut": ({"Nrad": "32", "Naz": "64", "ProfileCutoffOuter": "yes",
                "ProfileCutoffPointOuter": "2.0", "ProfileCutoffWidthOuter": "0.1",
                "ProfileCutoffInner": "yes", "ProfileCutoffPointInner": "0.6",
                "ProfileCutoffWidthInner": "0.05", "SigmaFloor": "1e-2",
                "ImposedDiskDrift": "1e-5", "FlaringIndex": "0.1"}, PLANET_1),
    # v_r = 0 (InitializeVradialZero), adiabatic
    "ic_vr0": (dict(ADIABATIC, Nrad="32", Naz="64", InitializeVradialZero="yes"), PLANET_1),
}


def make_ic_case(name, exe, addr):
    overrides, planets = IC_CASES[name]
    work = tempfile.mkdtemp(prefix=f"golden_{name}_")
    try:
        cfg = os.path.join(work, f"{name}.yml")
        write_config(cfg, overrides, planets)
        shutil.copy(cfg, os.path.join(HERE, "configs", f"{name}.yml"))
        dump = run_driver(exe, addr, cfg, os.path.join(work, "run"), "-", 0, 0, [])
        meta = read_meta(os.path.join(dump, "meta.txt"))
        nr, nphi = meta["Nr"], meta["Nphi"]
        c = dict(BASE)
        c.update(overrides)
        meta["case"] = name
        meta["sigma_slope"] = float(c["SigmaSlope"])
        meta["ic"] = {
            "imposed_disk_drift": float(c.get("ImposedDiskDrift", 0.0)),
            "vradial_zero": 1 if c.get("InitializeVradialZero", "no") == "yes" else 0,
            "cutoff_outer": 1 if c.get("ProfileCutoffOuter", "no") == "yes" else 0,
            "cutoff_outer_point": float(c.get("ProfileCutoffPointOuter", 1e300)),
            "cutoff_outer_width": float(c.get("ProfileCutoffWidthOuter", 1.0)),
            "cutoff_inner": 1 if c.get("ProfileCutoffInner", "no") == "yes" else 0,
            "cutoff_inner_point": float(c.get("ProfileCutoffPointInner", 0.0)),
            "cutoff_inner_width": float(c.get("ProfileCutoffWidthInner", 1.0)),
        }
        meta["bc_inner"] = c["InnerBoundary"].lower()
        meta["bc_outer"] = c["OuterBoundary"].lower()
        out = {"meta_json": np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)}
        shapes = {"sigma": (nr, nphi), "vrad": (nr + 1, nphi), "vazi": (nr, nphi),
                  "energy": (nr, nphi), "qplus": (nr, nphi), "qminus": (nr, nphi)}
        for f, shp in shapes.items():
            out[f"ic_{f}"] = load_bin(os.path.join(dump, f"state_ic_{f}.bin")).reshape(shp)
        for f in ("pressure", "soundspeed", "temperature", "viscosity"):
            out[f"ic_{f}"] = load_bin(os.path.join(dump, f"ic_{f}.bin")).reshape(nr, nphi)
        for f in ("sigma0", "vrad0", "vazi0", "energy0"):
            out["init_" + f] = load_bin(os.path.join(dump, f"init_{f}.bin")).reshape(shapes[f[:-1]])
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    finally:
        shutil.rmtree(work, ignore_errors=True)


def write_config(path: str, overrides: dict, planets) -> None:
    cfg = dict(BASE)
    cfg.update(overrides)
    lines = [f"{k}: {v}" for k, v in cfg.items() if k != "OutputDir"]
    lines.append(f"OutputDir: {cfg['OutputDir']}")
    lines.append("nbody:")
    lines += ["- name: Star", "  semi-major axis: 0.0 au", "  mass: 1.0 solMass",
              "  eccentricity: 0.0", "  radius: 1.0 solRadius", "  temperature: 0 K"]
    for k, pl in enumerate(planets):
        a, q = pl[0], pl[1]
        csf = pl[2] if len(pl) > 2 else 0.0
        ramp = pl[3] if len(pl) > 3 else 0.0
        lines += [f"- name: Planet{k}", f"  semi-major axis: {a} au", f"  mass: {q} solMass",
                  f"  cubic smoothing factor: {csf}", "  accretion efficiency: 0.0",
                  "  accretion method: none", "  eccentricity: 0", "  radius: 0.01 solRadius",
                  f"  ramp-up time: {ramp}"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def build() -> tuple[str, int]:
    subprocess.run(["make", "-f", "oracle/ref.mk", "-j8"], cwd=ROOT, check=True,
                   stdout=subprocess.DEVNULL)
    obj = os.path.join(REF_OUT, "refdump.o")
    exe = os.path.join(REF_OUT, "refdump")
    flags = ["-O2", "-fno-fast-math", "-ffp-contract=off", "-std=c++17", "-DNDEBUG",
             "-DOMPI_SKIP_MPICXX", "-DDISABLE_FFTW", "-DDISABLE_GSL",
             "-I/root/reference/src", "-I/opt/conda/include", "-fopenmp", "-no-pie", "-w"]
    subprocess.run(["g++", *flags, "-c", os.path.join(HERE, "refdump.cpp"), "-o", obj], check=True)
    mpilink = os.path.join(REF_OUT, "mpilib")
    subprocess.run(["g++", "-no-pie", "-o", exe, obj,
                    *[os.path.join(REF_OUT, x) for x in
                      ("libfargo_ref.a", "librebound.a", "libyamlcpp.a", "libunits.a")],
                    f"-L{mpilink}", f"-Wl,-rpath,{mpilink}", f"-Wl,-rpath-link,{mpilink}",
                    "-lmpi", "-lm", "-lrt", "-fopenmp"], check=True)
    nm = subprocess.run(["nm", exe], check=True, capture_output=True, text=True).stdout
    addr = [ln.split()[0] for ln in nm.splitlines() if ln.endswith(" _ZL6Nshift")]
    assert len(addr) == 1, "reference Nshift symbol not found"
    return exe, int(addr[0], 16)


def read_meta(path: str) -> dict:
    meta = {}
    for ln in open(path):
        if "=" in ln:
            k, v = ln.split("=", 1)
            v = v.strip()
            try:
                meta[k.strip()] = int(v)
            except ValueError:
                meta[k.strip()] = float(v)
    return meta


# cases whose final state is also written with the reference's snapshot
# writers (the .dat grids, misc.bin) next to its used_rad.dat and
# dimensions.dat
SNAPSHOT_CASES = {"c1_small", "c2_small"}


def run_driver(exe, addr, cfg, workdir, icdir, nsteps, stage_steps, checkpoints, simrun=False,
               snapshot=False, env_extra=None):
    out = os.path.join(workdir, "dump")
    os.makedirs(out, exist_ok=True)
    args = [exe, cfg, out, icdir, str(nsteps), str(stage_steps),
            ",".join(str(c) for c in checkpoints) or "0", hex(addr)]
    if simrun:
        args.append("simrun")
    env = dict(os.environ, OMP_NUM_THREADS="8")
    env.update(env_extra or {})
    if snapshot:
        env["REFDUMP_SNAPSHOT"] = "1"
    subprocess.run(args, cwd=workdir, check=True, stdout=subprocess.DEVNULL, env=env)
    return out


def load_bin(path, dtype=np.float64):
    return np.fromfile(path, dtype=dtype)


def case_inputs(name, exe, addr, work):
    """The config file of a case and the package's seeded initial condition
    for it (code units from a first run of the driver), written under work."""
    overrides, planets = CASES[name][:2]
    cfg = os.path.join(work, f"{name}.yml")
    write_config(cfg, overrides, planets)
    # pass 1: code units from the reference's own unit system
    probe = run_driver(exe, addr, cfg, os.path.join(work, "probe"), "-", 0, 0, [])
    meta = read_meta(os.path.join(probe, "meta.txt"))
    setup = DiskSetup(nr=meta["Nr"], nphi=meta["Nphi"], rmin=meta["RMIN"], rmax=meta["RMAX"],
                      sigma0=meta["sigma0"], sigma_slope=float(BASE["SigmaSlope"]),
                      aspect_ratio=meta["aspectratio"], flaring=meta["flaring"],
                      G=meta["G"], mstar=meta["hydro_center_mass"],
                      omega_frame=meta["OmegaFrame"], adiabatic=bool(meta["adiabatic"]),
                      gamma=meta["gamma"], noise_amp=1e-4, seed=1337)
    ic = initial_state(setup)
    icdir = os.path.join(work, "ic")
    os.makedirs(icdir)
    for f in ("sigma", "vrad", "vazi", "energy"):
        ic[f].tofile(os.path.join(icdir, f + ".bin"))
    return cfg, icdir


def make_case(name, exe, addr, keep_dir=None):
    overrides, planets, nsteps, stage_steps, checkpoints, full = CASES[name]
    work = tempfile.mkdtemp(prefix=f"golden_{name}_")
    try:
        cfg, icdir = case_inputs(name, exe, addr, work)
        shutil.copy(cfg, os.path.join(HERE, "configs", f"{name}.yml"))
        dump = run_driver(exe, addr, cfg, os.path.join(work, "run"), icdir, nsteps,
                          stage_steps, checkpoints, snapshot=name in SNAPSHOT_CASES)
        meta = read_meta(os.path.join(dump, "meta.txt"))
        meta["case"] = name
        meta["nsteps"] = nsteps
        meta["noise_amp"] = 1e-4
        meta["seed"] = 1337
        meta["sigma_slope"] = float(BASE["SigmaSlope"])
        meta["planets"] = [list(pl) for pl in planets]
        cfg = dict(BASE)
        cfg.update(overrides)
        meta["bc_inner"] = cfg["InnerBoundary"].lower()
        meta["bc_outer"] = cfg["OuterBoundary"].lower()
        out = {"meta_json": np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)}
        for g in ("Radii", "Rinf", "Rsup", "Rmed", "Surf", "InvSurf", "InvRmed", "InvRinf",
                  "InvDiffRmed", "InvDiffRsup", "InvDiffRsupRb", "TwoDiffRaSq"):
            out["geom_" + g] = load_bin(os.path.join(dump, f"geom_{g}.bin"))
        steps = load_bin(os.path.join(dump, "steps.bin")).reshape(nsteps, 16)
        out["steps"] = steps
        npl = meta["n_planets"]
        if meta.get("integrator", 0) == 1:
            # LeapFrog: the bodies of the first kick, then of the second
            pl = load_bin(os.path.join(dump, "planets.bin")).reshape(nsteps, 2, npl, 4)
            out["planets"] = pl[:, 0]
            out["planets_mid"] = pl[:, 1]
        else:
            out["planets"] = load_bin(os.path.join(dump, "planets.bin")).reshape(nsteps, npl, 4)
        out["nshift"] = load_bin(os.path.join(dump, "nshift.bin"), np.int32).reshape(nsteps, nr)
        # the disk's gravity on every body after each step (x, y, r, Klahr radius, ax, ay)
        out["diskacc"] = load_bin(os.path.join(dump, "diskacc.bin")).reshape(nsteps, npl, 6)
        # the reference's gas totals after each step, one thread (mass, angular
        # momentum; the same up to the radius in column 4)
        out["totals"] = load_bin(os.path.join(dump, "totals.bin")).reshape(nsteps, 5)
        shapes = {"sigma": (nr, nphi), "vrad": (nr + 1, nphi), "vazi": (nr, nphi),
                  "energy": (nr, nphi), "qplus": (nr, nphi), "qminus": (nr, nphi)}
        for f in ("sigma0", "vrad0", "vazi0", "energy0"):
            shp = shapes[f[:-1]]
            arr = load_bin(os.path.join(dump, f"init_{f}.bin")).reshape(shp)
            if full:
                out["init_" + f] = arr
            else:
                # reference values equal the initial condition: store a checksum only
                out["init_" + f + "_ringsum"] = arr.sum(axis=1)
        for tag in ["0"] + [str(c) for c in checkpoints]:
            # whole arrays for the start and the final checkpoint of small grids,
            # per-ring moments and sample rings otherwise
            whole = full and (tag == "0" or tag == str(checkpoints[-1]) or nr * nphi <= 8192)
            for f, shp in shapes.items():
                arr = load_bin(os.path.join(dump, f"state_{tag}_{f}.bin")).reshape(shp)
                if whole:
                    out[f"state_{tag}_{f}"] = arr
                else:
                    out[f"state_{tag}_{f}_ringsum"] = arr.sum(axis=1)
                    out[f"state_{tag}_{f}_ringmin"] = arr.min(axis=1)
                    out[f"state_{tag}_{f}_ringmax"] = arr.max(axis=1)
                    rows = np.array([1, 2, shp[0] // 4, shp[0] // 2, 3 * shp[0] // 4,
                                     shp[0] - 3, shp[0] - 2])
                    out[f"state_{tag}_{f}_rows"] = arr[rows]
                    out[f"state_{tag}_{f}_rowidx"] = rows
        if stage_steps:
            names = sorted(fn for fn in os.listdir(dump) if fn.startswith("stage_"))
            fields = ["sigma", "vrad", "vazi", "energy", "pressure", "soundspeed",
                      "scaleheight", "viscosity", "potential", "qr", "qphi", "divv",
                      "taurr", "taurp", "taupp", "qplus", "qminus", "temperature"]
            vec = {"vrad", "taurp"}
            # keep the first stage whole, then only the fields a stage changed
            prev = {}
            for fn in names:
                key = fn[:-4]
                fld = [f for f in fields if key.endswith("_" + f)]
                fld = max(fld, key=len)
                shp = (nr + 1, nphi) if fld in vec else (nr, nphi)
                arr = load_bin(os.path.join(dump, fn)).reshape(shp)
                if fld not in prev or not np.array_equal(prev[fld], arr):
                    out[key] = arr
                prev[fld] = arr
        if name in SNAPSHOT_CASES:
            # file bytes, as written by the reference
            snap = os.path.join(dump, "snapshot")
            for fn in sorted(os.listdir(snap)):
                out["snap_" + fn] = np.fromfile(os.path.join(snap, fn), dtype=np.uint8)
            outdir = os.path.join(work, "run", BASE["OutputDir"])
            for fn in ("used_rad.dat", "dimensions.dat"):
                out["grid_" + fn] = np.fromfile(os.path.join(outdir, fn), dtype=np.uint8)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        if keep_dir:
            shutil.copytree(dump, keep_dir, dirs_exist_ok=True)
        return out
    finally:
        shutil.rmtree(work, ignore_errors=True)


def check_simrun(exe, addr, name="c1_small"):
    """The driver's step loop must equal the reference's sim::run bit for bit."""
    overrides, planets, nsteps, _, _, _ = CASES[name]
    data = np.load(os.path.join(HERE, f"{name}.npz"))
    meta = json.loads(bytes(data["meta_json"]).decode())
    work = tempfile.mkdtemp(prefix="golden_simrun_")
    try:
        cfg = os.path.join(work, "cfg.yml")
        write_config(cfg, overrides, planets)
        setup = DiskSetup(nr=meta["Nr"], nphi=meta["Nphi"], rmin=meta["RMIN"], rmax=meta["RMAX"],
                          sigma0=meta["sigma0"], sigma_slope=meta["sigma_slope"],
                          aspect_ratio=meta["aspectratio"], flaring=meta["flaring"], G=meta["G"],
                          mstar=meta["hydro_center_mass"], omega_frame=meta["OmegaFrame"],
                          adiabatic=bool(meta["adiabatic"]), gamma=meta["gamma"])
        ic = initial_state(setup)
        icdir = os.path.join(work, "ic")
        os.makedirs(icdir)
        for f in ("sigma", "vrad", "vazi", "energy"):
            ic[f].tofile(os.path.join(icdir, f + ".bin"))
        dump = run_driver(exe, addr, cfg, os.path.join(work, "run"), icdir, nsteps, 0, [],
                          simrun=True)
        for f in ("sigma", "vrad", "vazi", "energy"):
            a = load_bin(os.path.join(dump, f"state_simrun_{f}.bin"))
            b = data[f"state_{nsteps}_{f}"].ravel()
            assert np.array_equal(a, b), f"{name}: sim::run differs from driver loop in {f}"
        print(f"{name}: driver loop == sim::run (bit-exact)")
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main(argv):
    os.makedirs(os.path.join(HERE, "configs"), exist_ok=True)
    exe, addr = build()
    names = argv or list(CASES) + list(IC_CASES)
    for name in names:
        if name in IC_CASES:
            make_ic_case(name, exe, addr)
        else:
            make_case(name, exe, addr)
        print("wrote", name)
    for name in ("c1_small", "lf_iso", "lf_adi"):
        if name in names:
            check_simrun(exe, addr, name)


if __name__ == "__main__":
    main(sys.argv[1:])
