"""Grids of awkward sizes against the oracle, step by step in FP64.

Ring counts that leave partial tiles in the ring walks and the radial sweep
(odd and prime Nr) and sector counts that are not a multiple of the azimuthal
sweep's 120-sector segment (a prime below one segment, one sector past a
segment, a prime past two), on the large adiabatic disk's physics (beta
cooling, three planets, von Neumann-Richtmyer artificial viscosity) and on the
locally isothermal disk's, with the Euler and the LeapFrog integrators.
At every step the FARGO shifts must be identical and the clock within 1e-12;
after the run every evolved field must be within 1e-11 relative (max |d| /
max |field|) of the oracle's.  Split storage on two of the grids is held to
the FP64 run of the same grid.
"""
import dataclasses

import numpy as np
import pytest

from fixtures import rel_err

pytestmark = pytest.mark.gpu

STEPS = 40
# (workload, Nr, Nphi, integrator): 0 Euler, 1 LeapFrog
SIZES = [("adi_beta_4096x12288", 37, 67, 0), ("adi_beta_4096x12288", 61, 131, 0),
         ("adi_beta_4096x12288", 97, 241, 0), ("iso_128x384", 53, 127, 0), ("iso_128x384", 29, 121, 0),
         ("adi_beta_4096x12288", 43, 139, 1), ("iso_128x384", 31, 97, 1)]
SPLIT_TOL = 2e-5    # split storage against FP64 on the same ragged grid, relative


def _config(name, nr, nphi, integrator=0, precision=0):
    from fargocpt_amd import workloads as W
    from fargocpt_amd.ic import log_radii
    cfg = W.disk_config(name, precision)
    cfg.update({"nr": nr, "nphi": nphi, "radii": log_radii(nr, cfg["rmin"], cfg["rmax"]),
                "integrator": integrator})
    return cfg


def _start(disk, name, nr, nphi):
    """The workload's start-up sequence on the resized grid."""
    from fargocpt_amd import workloads as W
    from fargocpt_amd.ic import initial_state
    ic = initial_state(dataclasses.replace(W.setup(name), nr=nr, nphi=nphi))
    disk.upload_reference(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    disk.upload_state(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    disk.set_time(0.0, 0.1, 0)
    disk.calculate_timestep()
    disk.apply_boundary_condition(0.0, False)
    disk.calculate_timestep()
    bodies, ind = W.planet_inputs(name)
    disk.set_step_params(bodies, ind)


@pytest.mark.parametrize("name,nr,nphi,integrator", SIZES)
def test_ragged_grid_follows_oracle(name, nr, nphi, integrator):
    from fargocpt_amd.capi import Disk
    from oracle import OracleDisk
    print(f"{name} resized to {nr} x {nphi}, integrator {integrator}: {STEPS} steps against the oracle",
          flush=True)
    cfg = _config(name, nr, nphi, integrator)
    gpu = Disk(cfg)
    cpu = OracleDisk(cfg)
    for d in (gpu, cpu):
        _start(d, name, nr, nphi)
    for s in range(STEPS):
        gpu.run_steps(1, True)
        gpu.synchronize()
        cpu.run_step()
        assert np.array_equal(gpu.nshift(), cpu.nshift()), (nr, nphi, s)
        assert abs(gpu.get_time()[0] - cpu.get_time()[0]) <= 1e-12 * cpu.get_time()[0], (nr, nphi, s)
    fields = ("SIGMA", "VRAD", "VAZI") + (("ENERGY",) if cfg["adiabatic"] else ())
    errs = {f: rel_err(gpu.download_field(f), cpu.download_field(f)) for f in fields}
    print(f"errors {errs}", flush=True)
    gpu.close()
    cpu.close()
    bad = {f: e for f, e in errs.items() if not e <= 1e-11}
    assert not bad, bad


@pytest.mark.parametrize("name,nr,nphi", [("adi_beta_4096x12288", 61, 131), ("iso_128x384", 53, 127)])
def test_ragged_grid_split_storage(name, nr, nphi):
    """Split storage on a ragged grid against FP64 storage, same start and
    steps: the FARGO shifts identical but at most at one step (Nshift =
    floor(N + 0.5) may flip where N lies within the split rounding of a
    half-integer), every field within SPLIT_TOL."""
    from fargocpt_amd.capi import Disk
    print(f"{name} resized to {nr} x {nphi}: split against FP64, {STEPS} steps", flush=True)
    disks = []
    for prec in (0, 1):
        d = Disk(_config(name, nr, nphi, 0, prec))
        _start(d, name, nr, nphi)
        disks.append(d)
    flips = 0
    for _ in range(STEPS):
        for d in disks:
            d.run_steps(1, True)
            d.synchronize()
        flips += not np.array_equal(disks[0].nshift(), disks[1].nshift())
    assert flips <= 1, flips
    fp64, split = disks
    fields = ("SIGMA", "VRAD", "VAZI") + (("ENERGY",) if "adi" in name else ())
    errs = {f: rel_err(split.download_field(f), fp64.download_field(f)) for f in fields}
    print(f"errors {errs}", flush=True)
    for d in disks:
        d.close()
    bad = {f: e for f, e in errs.items() if not e <= SPLIT_TOL}
    assert not bad, bad
