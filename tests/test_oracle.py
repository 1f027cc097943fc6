"""The CPU oracle against the reference FargoCPT fixtures and against closed
forms.  The oracle is a C restatement of the reference's loops; built without
fast-math and FMA contraction it must reproduce the reference bit for bit."""
import numpy as np
import pytest

from fixtures import (check_stages, compare_state, disk_config, initial_condition, load_case,
                      run_case, totals_bound, totals_magnitudes)
from oracle import OracleDisk

SMALL = ["c1_small", "c2_small", "c5_small", "mix_small", "none_small", "smooth_small",
         "nostar_small"]


@pytest.mark.parametrize("case", ["c1_small", "c1_iso", "c5_small"])
def test_geometry_bit_exact(case, oracle_lib):
    meta, data = load_case(case)
    d = OracleDisk(disk_config(meta))
    nr = meta["Nr"]
    for g, n in [("Radii", nr + 1), ("Rinf", nr + 1), ("Rsup", nr), ("Rmed", nr + 1),
                 ("Surf", nr), ("InvSurf", nr), ("InvRmed", nr + 1), ("InvRinf", nr + 1),
                 ("InvDiffRmed", nr + 1), ("InvDiffRsup", nr), ("InvDiffRsupRb", nr),
                 ("TwoDiffRaSq", nr)]:
        assert np.array_equal(d.geom(g, n), data["geom_" + g]), g
    ids = [meta[k] for k in ("damp_inner_scalar", "damp_inner_vector", "damp_outer_scalar",
                             "damp_outer_vector")]
    assert d.damping_ids() == ids


@pytest.mark.parametrize("case", SMALL)
@pytest.mark.parametrize("step", [1, 2])
def test_stages_bit_exact(case, step, oracle_lib):
    meta, data = load_case(case)
    if f"stage_{step}_00_start_sigma" not in data.files and step > 1:
        if not any(k.startswith(f"stage_{step}_") for k in data.files):
            pytest.skip("no dump of this step")
    d = OracleDisk(disk_config(meta))
    errors, nshift_ok = check_stages(d, meta, data, step)
    bad = {k: v for k, v in errors.items() if v != 0.0}
    assert not bad, bad
    assert nshift_ok


@pytest.mark.parametrize("case", SMALL + ["c1_iso", "lf_iso", "lf_adi"])
def test_full_run_bit_exact(case, oracle_lib):
    meta, data = load_case(case)
    d = OracleDisk(disk_config(meta))
    worst_dt, bad_shift, res = run_case(d, meta, data)
    assert worst_dt == 0.0
    assert bad_shift == 0
    for tag, errs in res.items():
        assert all(v == 0.0 for v in errs.values()), (tag, errs)


def test_adiabatic_512x1536_500_steps(oracle_lib):
    """The full adiabatic configuration: per-ring moments and sample rings of
    all fields after 500 steps, every dt and every FARGO shift."""
    meta, data = load_case("c2_adi")
    d = OracleDisk(disk_config(meta))
    worst_dt, bad_shift, res = run_case(d, meta, data)
    assert worst_dt == 0.0 and bad_shift == 0
    for tag, errs in res.items():
        assert all(v == 0.0 for v in errs.values()), (tag, errs)


# --- closed-form checks that do not depend on the reference fixtures --------

def _small_disk(nr=24, nphi=48, adiabatic=False):
    meta, _ = load_case("c2_small" if adiabatic else "c1_small")
    meta = dict(meta, Nr=nr, Nphi=nphi)
    return meta, OracleDisk(disk_config(meta))


def test_pure_fargo_shift_is_a_rotation(oracle_lib):
    """v_r = 0 and v_phi = N_i Rmed_i dphi / dt make the residual and the
    uniform residual velocities vanish (up to the rounding of N_i Rmed_i dphi
    / dt): transport is then np.roll by N_i, and the shifts are exact."""
    meta, d = _small_disk()
    nr, ns = meta["Nr"], meta["Nphi"]
    rng = np.random.default_rng(5)
    sigma = 1.0 + rng.random((nr, ns))
    rmed = d.geom("Rmed", nr)
    dt = 0.01
    dphi = 2 * np.pi / ns
    shifts = (np.arange(nr) % 7) - 3
    vazi = np.repeat((shifts * rmed * dphi / dt)[:, None], ns, axis=1)
    d.upload_state(sigma, np.zeros((nr + 1, ns)), vazi, np.zeros((nr, ns)))
    d.transport(dt)
    got = d.download_field("SIGMA")
    want = np.stack([np.roll(sigma[i], shifts[i]) for i in range(nr)])
    assert np.array_equal(d.nshift(), shifts)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


def test_transport_conserves_mass(oracle_lib):
    """With v_r = 0 on the boundary interfaces the flux form conserves
    sum(Surf * Sigma) to rounding (TransportEuler.cpp:587-597, 652-661)."""
    meta, d = _small_disk(adiabatic=True)
    nr, ns = meta["Nr"], meta["Nphi"]
    rng = np.random.default_rng(7)
    sigma = 1.0 + 0.3 * rng.random((nr, ns))
    vrad = 1e-3 * rng.standard_normal((nr + 1, ns))
    vrad[[0, 1, nr - 1, nr]] = 0.0
    rmed = d.geom("Rmed", nr)
    vazi = (rmed ** -0.5)[:, None] - rmed[:, None] + 1e-3 * rng.standard_normal((nr, ns))
    energy = sigma * 1e-3
    d.upload_state(sigma, vrad, vazi, energy)
    surf = d.geom("Surf", nr)
    m0 = np.sum(surf[:, None] * sigma)
    d.transport(0.01)
    m1 = np.sum(surf[:, None] * d.download_field("SIGMA"))
    assert abs(m1 - m0) / m0 < 1e-14


def test_potential_closed_form(oracle_lib):
    meta, d = _small_disk()
    nr, ns = meta["Nr"], meta["Nphi"]
    ic = initial_condition(dict(meta))
    d.upload_state(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    planets = [(0.0, 0.0, 1.0, 0.0), (0.8, 0.3, 1e-3, 0.0)]
    d.set_step_params(planets, (1e-4, -2e-4))
    d.nbody_potential()
    rmed = d.geom("Rmed", nr)
    phi = 2 * np.pi * np.arange(ns) / ns
    x = rmed[:, None] * np.cos(phi)[None, :]
    y = rmed[:, None] * np.sin(phi)[None, :]
    H = d.download_field("SCALE_HEIGHT")
    want = np.zeros((nr, ns))
    for (px, py, m, _) in planets:
        want += -meta["G"] * m / np.sqrt((x - px) ** 2 + (y - py) ** 2 + (0.6 * H) ** 2)
    want += -1e-4 * x + 2e-4 * y
    np.testing.assert_allclose(d.download_field("POTENTIAL"), want, rtol=1e-13, atol=0)


TOTALS_CASES = ["c1_small", "c2_small", "c5_small", "mix_small", "none_small", "c1_iso",
                "lf_iso", "lf_adi", "smooth_small", "nostar_small"]


@pytest.mark.parametrize("case", TOTALS_CASES)
def test_totals_match_reference(case, oracle_lib):
    """fo_totals against the reference's own quantities::gas_total_mass and
    gas_angular_momentum (one thread, loop order) at every checkpoint the
    fixture holds whole, over the whole disk and up to the middle ring's
    radius.  Bound: the compensated sum is within 2u of the exact sum and the
    reference's sequential sum within (n - 1) u sum |term| of it."""
    meta, data = load_case(case)
    d = OracleDisk(disk_config(meta))
    nr = meta["Nr"]
    geom = {g: data["geom_" + g] for g in ("Surf", "Rmed")}
    tags = [k[len("state_"):-len("_sigma")] for k in data.files
            if k.startswith("state_") and k.endswith("_sigma") and k[6:-6].isdigit()]
    checked = 0
    for tag in tags:
        step = int(tag)
        if step == 0:
            continue
        sig, vaz = data[f"state_{tag}_sigma"], data[f"state_{tag}_vazi"]
        d.upload_state(sig, data[f"state_{tag}_vrad"], vaz, None)
        ref = data["totals"][step - 1]
        for radius, (m_ref, l_ref) in ((1e300, ref[0:2]), (ref[4], ref[2:4])):
            m, l = d.totals(radius)
            am, al, n = totals_magnitudes(sig, vaz, geom["Surf"], geom["Rmed"], meta["OmegaFrame"],
                                          radius, nr)
            assert abs(m - m_ref) <= totals_bound(n, am), (tag, radius, m, m_ref)
            assert abs(l - l_ref) <= totals_bound(n, al), (tag, radius, l, l_ref)
            checked += 1
    assert checked >= 2


def test_totals_closed_forms(oracle_lib):
    """Closed forms that a dropped ring, a wrong frame term or a wrong
    azimuthal average fail: uniform Sigma gives M = Sigma pi (Rsup[nr-2]^2 -
    Rinf[1]^2) (the ring areas telescope); v_phi = -Omega_F R gives L = 0
    exactly; rigid rotation v_phi + Omega_F R = w R with uniform Sigma gives
    L = Sigma w sum_i Nphi Surf_i Rmed_i^2; the radius cut includes the ring
    whose Rmed equals it."""
    meta, data = load_case("c1_small")
    d = OracleDisk(disk_config(meta))
    nr, ns = meta["Nr"], meta["Nphi"]
    rinf, rsup = data["geom_Rinf"], data["geom_Rsup"]
    rmed, surf = data["geom_Rmed"][:nr], data["geom_Surf"]
    om = meta["OmegaFrame"]
    s0 = 3.0e-5
    sig = np.full((nr, ns), s0)
    vrad = np.zeros((nr + 1, ns))
    vaz = np.repeat((-(om * rmed))[:, None], ns, axis=1)
    d.upload_state(sig, vrad, vaz, None)
    m, l = d.totals(1e300)
    m_closed = s0 * np.pi * (rsup[nr - 2] ** 2 - rinf[1] ** 2)
    assert abs(m - m_closed) <= 1e-13 * m_closed
    assert l == 0.0
    w = 0.7
    vaz = np.repeat((w * rmed - om * rmed)[:, None], ns, axis=1)
    d.upload_state(sig, vrad, vaz, None)
    m, l = d.totals(1e300)
    l_closed = s0 * w * float(np.sum(ns * surf[1:nr - 1] * rmed[1:nr - 1] ** 2))
    assert abs(l - l_closed) <= 1e-13 * l_closed
    k = nr // 2
    m_in, _ = d.totals(rmed[k])
    m_out, _ = d.totals(np.nextafter(rmed[k], 0.0))
    assert abs((m_in - m_out) - s0 * ns * surf[k]) <= 1e-12 * m_in


def _force_flags(meta):
    from fargocpt_amd.capi import (FORCE_CORRECT_SELFGRAVITY, FORCE_NO_STAR_SMOOTHING,
                                   FORCE_SMOOTHING_PLANETLOC)
    return ((FORCE_CORRECT_SELFGRAVITY if meta.get("correct_disk_selfgravity", 1) else 0) |
            (FORCE_NO_STAR_SMOOTHING if meta.get("compatibility_no_star_smoothing", 0) else 0) |
            (FORCE_SMOOTHING_PLANETLOC if meta.get("compatibility_smoothing_planetloc", 0) else 0))


@pytest.mark.parametrize("case", ["c1_small", "c5_small", "mix_small", "smooth_small", "nostar_small"])
def test_disk_force_bit_exact(case, oracle_lib):
    """ComputeDiskOnPlanetAccel after every step of the reference run (one
    OpenMP thread, loop order), recomputed by the oracle from its own replay."""
    from fixtures import start_run, step_inputs
    meta, data = load_case(case)
    d = OracleDisk(disk_config(meta))
    start_run(d, meta)
    acc = data["diskacc"]
    flags = _force_flags(meta)
    for s in range(1, meta["nsteps"] + 1):
        planets, ind = step_inputs(data, s)
        d.set_step_params(planets, ind)
        d.run_step()
        bodies = [tuple(b[:4]) for b in acc[s - 1]]
        got = d.disk_on_planet_accel(bodies, flags)
        assert np.array_equal(got, acc[s - 1][:, 4:6]), (s, got, acc[s - 1][:, 4:6])


def test_disk_force_closed_forms(oracle_lib):
    """One massive cell: the point-mass acceleration G m d / (d^2 + eps^2)^1.5;
    a ring-uniform disk with the self-gravity correction (Force.cpp:64
    subtracts the ring mean): exactly zero."""
    from fargocpt_amd.capi import FORCE_CORRECT_SELFGRAVITY
    meta, d = _small_disk()
    nr, ns = meta["Nr"], meta["Nphi"]
    ic = initial_condition(dict(meta))
    sigma = np.zeros((nr, ns))
    i0, j0 = nr // 2, ns // 3
    sigma[i0, j0] = 2.5e-4
    d.upload_state(sigma, ic["vrad"], ic["vazi"], ic["energy"])
    H = d.download_field("SCALE_HEIGHT")
    surf = d.geom("Surf", nr)
    rmed = d.geom("Rmed", nr)
    phi = 2 * np.pi * j0 / ns
    xc, yc = rmed[i0] * np.cos(phi), rmed[i0] * np.sin(phi)
    body = (0.9, -0.4, float(np.hypot(0.9, -0.4)), 0.0)
    got = d.disk_on_planet_accel([body], 0)[0]
    eps = 0.6 * H[i0, j0]
    dx, dy = xc - body[0], yc - body[1]
    r3 = (dx * dx + dy * dy + eps * eps) ** 1.5
    m = surf[i0] * 2.5e-4
    np.testing.assert_allclose(got, [meta["G"] * m * dx / r3, meta["G"] * m * dy / r3], rtol=1e-13)
    # sign: the disk pulls the body towards the cell
    assert got[0] * dx > 0 and got[1] * dy > 0
    # ring values k * 2^-17: every partial sum of the ring mean is exact
    ring = np.repeat((1.0 + np.arange(nr))[:, None] * 2.0 ** -17, ns, axis=1)
    d.upload_state(ring, ic["vrad"], ic["vazi"], ic["energy"])
    zero = d.disk_on_planet_accel([(0.0, 0.0, 0.0, 0.0), body], FORCE_CORRECT_SELFGRAVITY)
    assert np.all(zero == 0.0)


@pytest.mark.parametrize("case", ["ic_iso", "ic_adi", "ic_cut", "ic_vr0"])
def test_initial_condition_bit_exact(case, oracle_lib):
    """init_physics for profile setups: Sigma with cutoffs and floor, E with
    the temperature floor, the derived fields, v_phi and the numerical
    viscous v_r, the boundary call and the damping references -- the
    reference's own values after init_physics, bit for bit."""
    from fixtures import IC_FIELDS, IC_REFERENCE, ic_params
    meta, data = load_case(case)
    d = OracleDisk(disk_config(meta))
    d.init_profiles(**ic_params(meta))
    for key, name in IC_FIELDS:
        if not meta["adiabatic"] and key == "energy":
            continue
        got = d.download_field(name)
        assert np.array_equal(got, data["ic_" + key]), (key, np.max(np.abs(got - data["ic_" + key])))
    for key, name in IC_REFERENCE:
        if not meta["adiabatic"] and key == "energy0":
            continue
        assert np.array_equal(d.download_field(name), data["init_" + key]), key
