"""Disk-on-planet gravity (ComputeDiskOnPlanetAccel, Force.cpp:23-122) on
the GPU against the oracle, through the C-ABI.

The oracle sums the cells in the reference's loop order; the GPU reduces per
block and then over blocks, so the two differ by summation order only.  The
tolerance is written against the sum of the absolute contributions
(|G m dx / d^3| over every cell), the scale on which reordering errors live:
1e-12 of it in FP64 storage.  With the self-gravity correction each cell
contributes its excess over the ring mean, and the two ring means (sums of
N_phi values in different orders) may differ by up to N_phi * 2^-52 of the
mean; that difference enters every cell of the ring alike, so it is bounded
separately by N_phi * 2^-52 * sum |G m_mean dx / d^3|.  In split storage the density is the decoded
background + FP32 deviation (the same numbers the oracle receives), but the
scale height of an adiabatic disk is read from split storage, so the bound is
2e-7 of the scale (2^-24 relative error of H, twice, through d^3).
"""
import numpy as np
import pytest

from fixtures import disk_config, load_case, run_case

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
SPLIT_TOL = 2e-7


def _flags(meta):
    from fargocpt_amd.capi import (FORCE_CORRECT_SELFGRAVITY, FORCE_NO_STAR_SMOOTHING,
                                   FORCE_SMOOTHING_PLANETLOC)
    return ((FORCE_CORRECT_SELFGRAVITY if meta.get("correct_disk_selfgravity", 1) else 0) |
            (FORCE_NO_STAR_SMOOTHING if meta.get("compatibility_no_star_smoothing", 0) else 0) |
            (FORCE_SMOOTHING_PLANETLOC if meta.get("compatibility_smoothing_planetloc", 0) else 0))


def _abs_scale(d, bodies, flags, G, thickness):
    """Per body and component: (sum over cells of |G m dx / d_sm^3|, with m
    the summed mass -- the cell's, or its excess over the ring mean when the
    self-gravity correction is on; the bound on the ring-mean rounding)."""
    from fargocpt_amd.capi import FORCE_CORRECT_SELFGRAVITY
    sigma = d.download_field("SIGMA")
    H = d.download_field("SCALE_HEIGHT")
    nr, ns = sigma.shape
    surf = d.geom("Surf", nr)[:, None]
    rmed = d.geom("Rmed", nr)[:, None]
    phi = 2 * np.pi * np.arange(ns)[None, :] / ns
    x, y = rmed * np.cos(phi), rmed * np.sin(phi)
    corrected = bool(flags & FORCE_CORRECT_SELFGRAVITY)
    mean = sigma.mean(axis=1, keepdims=True)
    m = np.abs(sigma - mean) * surf if corrected else np.abs(sigma) * surf
    m_mean = np.abs(mean) * surf * (ns * 2.0 ** -52 if corrected else 0.0)
    scale, mean_err = [], []
    for b in bodies:
        dx, dy = np.abs(x - b[0]), np.abs(y - b[1])
        d3 = (dx * dx + dy * dy + (thickness * H) ** 2) ** 1.5
        scale.append([G * np.sum(m * dx / d3), G * np.sum(m * dy / d3)])
        mean_err.append([G * np.sum(m_mean * dx / d3), G * np.sum(m_mean * dy / d3)])
    return np.array(scale), np.array(mean_err)


def _check_against_oracle(gpu, cfg, bodies, flags, G, thickness, tol, tag):
    from oracle import OracleDisk
    cfg64 = dict(cfg)
    cfg64["precision"] = 0
    o = OracleDisk(cfg64)
    fields = [gpu.download_field(f) for f in ("SIGMA", "VRAD", "VAZI")]
    energy = gpu.download_field("ENERGY") if cfg.get("adiabatic") else None
    o.upload_state(*fields, energy)
    want = o.disk_on_planet_accel(bodies, flags)
    got = gpu.disk_on_planet_accel(bodies, flags)
    scale, mean_err = _abs_scale(o, bodies, flags, G, thickness)
    o.close()
    err = np.abs(got - want)
    print(f"{tag}: max |err| / scale {np.max(err / scale):.2e} "
          f"max |err| / bound {np.max(err / (tol * scale + mean_err)):.2e}", flush=True)
    assert np.all(np.isfinite(got)), tag
    assert np.all(err <= tol * scale + mean_err), (tag, got, want, scale, mean_err)
    return got, want


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", ["c1_small", "c5_small", "mix_small", "c2_adi"])
def test_force_fixture_final_state(case, precision):
    from fargocpt_amd.capi import Disk
    meta, data = load_case(case)
    print(f"force case={case} precision={precision}", flush=True)
    cfg = dict(disk_config(meta))
    cfg["precision"] = precision
    d = Disk(cfg)
    run_case(d, meta, data)
    if "diskacc" in data.files:
        bodies = [tuple(b[:4]) for b in data["diskacc"][-1]]
    else:
        bodies = [(0.0, 0.0, 0.0, 0.0)] + [(p[0], p[1], float(np.hypot(p[0], p[1])), 0.0)
                                           for p in data["planets"][-1][1:]]
    # add a body inside the disk, closer to the cells than any planet
    bodies.append((0.31, 0.52, float(np.hypot(0.31, 0.52)), 0.0))
    _check_against_oracle(d, cfg, bodies, _flags(meta), meta["G"], meta["thickness_smoothing"],
                          FP64_TOL if precision == 0 else SPLIT_TOL, case)
    # the same bodies without the self-gravity correction
    _check_against_oracle(d, cfg, bodies, 0, meta["G"], meta["thickness_smoothing"],
                          FP64_TOL if precision == 0 else SPLIT_TOL, case + " uncorrected")
    d.close()


def test_force_smoothing_variants():
    """Star without smoothing and planet-location smoothing length
    (the two compatibility switches of Force.cpp:44-59)."""
    from fargocpt_amd.capi import (Disk, FORCE_CORRECT_SELFGRAVITY, FORCE_NO_STAR_SMOOTHING,
                                   FORCE_SMOOTHING_PLANETLOC)
    meta, data = load_case("c1_small")
    cfg = dict(disk_config(meta))
    d = Disk(cfg)
    run_case(d, meta, data, nsteps=5)
    bodies = [(0.0, 0.0, 0.0, 0.0), (0.98, 0.2, float(np.hypot(0.98, 0.2)), 0.05)]
    for flags in (FORCE_NO_STAR_SMOOTHING, FORCE_SMOOTHING_PLANETLOC,
                  FORCE_NO_STAR_SMOOTHING | FORCE_SMOOTHING_PLANETLOC | FORCE_CORRECT_SELFGRAVITY):
        _check_against_oracle(d, cfg, bodies, flags, meta["G"], meta["thickness_smoothing"],
                              1e-10, f"flags={flags}")
    d.close()


def test_force_full_size_2048x6144():
    """The 2048 x 6144 isothermal disk after 20 steps: star, planet and six
    more bodies (eight, the kernel's maximum per call)."""
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Disk, FORCE_CORRECT_SELFGRAVITY
    name = "iso_2048x6144"
    print(f"force {name}", flush=True)
    cfg = W.disk_config(name)
    d = Disk(cfg)
    W.start(d, name)
    d.run_steps(20, True)
    d.synchronize()
    rng = np.random.default_rng(3)
    bodies = [(0.0, 0.0, 0.0, 0.0), (1.0, 0.0, 1.0, 0.0)]
    for _ in range(6):
        r, ph = rng.uniform(0.5, 2.0), rng.uniform(0, 2 * np.pi)
        bodies.append((r * np.cos(ph), r * np.sin(ph), r, rng.uniform(0.0, 0.1)))
    setup = W.setup(name)
    _check_against_oracle(d, cfg, bodies, FORCE_CORRECT_SELFGRAVITY, setup.G,
                          cfg.get("thickness_smoothing", 0.6), FP64_TOL, name)
    d.close()
