"""The fused step kernels (source terms, artificial viscosity, viscous stress
+ update, one pass each) against the separate kernels of the per-function
entry points, through the C-ABI.

In FP64 storage the fused passes evaluate the same per-cell expressions on
the same values, so every evolved and derived field must come out identical,
bit for bit, after whole fixture runs (TW and SN artificial viscosity, none,
isothermal and adiabatic, beta cooling, damping, outflow boundaries, Euler
and LeapFrog).  In
split storage the separate kernels round the intermediate tensors (Q, tau) to
FP32 and the fused ones keep them in FP64, so the two agree to the split
storage's own tolerance.  The batched step (fused) is checked against single
disks elsewhere.
"""
import numpy as np
import pytest

from fixtures import disk_config, load_case, rel_err, run_case

pytestmark = pytest.mark.gpu

FIELDS = ["SIGMA", "VRAD", "VAZI", "ENERGY", "PRESSURE", "SOUNDSPEED", "SCALE_HEIGHT",
          "VISCOSITY", "TEMPERATURE", "QPLUS", "QMINUS"]
CASES = ["c1_small", "c2_small", "c5_small", "mix_small", "none_small", "c1_iso", "lf_iso",
         "lf_adi"]


def _run(meta, data, precision, fused, nsteps):
    from fargocpt_amd.capi import Disk
    cfg = dict(disk_config(meta))
    cfg["precision"] = precision
    d = Disk(cfg)
    d.set_fused_step(fused)
    run_case(d, meta, data, nsteps=nsteps)
    out = {f: d.download_field(f) for f in FIELDS}
    shifts = d.nshift()
    d.close()
    return out, shifts


@pytest.mark.parametrize("case", CASES)
def test_fused_step_bit_identical_fp64(case):
    meta, data = load_case(case)
    n = min(meta["nsteps"], 20)
    print(f"fused vs separate, FP64, case={case} steps={n}", flush=True)
    a, sa = _run(meta, data, 0, True, n)
    b, sb = _run(meta, data, 0, False, n)
    assert np.array_equal(sa, sb)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), (f, rel_err(a[f], b[f]))


@pytest.mark.parametrize("case", ["c2_small", "mix_small", "c1_iso"])
def test_fused_step_split(case):
    meta, data = load_case(case)
    n = min(meta["nsteps"], 20)
    print(f"fused vs separate, split, case={case} steps={n}", flush=True)
    a, _ = _run(meta, data, 1, True, n)
    b, _ = _run(meta, data, 1, False, n)
    for f in ("SIGMA", "VRAD", "VAZI", "ENERGY"):
        assert rel_err(a[f], b[f]) <= 2e-6, (f, rel_err(a[f], b[f]))


STEP_FIELDS = ("SIGMA", "VRAD", "VAZI", "ENERGY", "TEMPERATURE", "QPLUS", "QMINUS")
TENSOR = ("TAU_R_R", "TAU_R_PHI", "TAU_PHI_PHI", "DIV_V")


def _check(got, want, precision, what):
    for f in want:
        if precision == 0:
            assert np.array_equal(got[f], want[f]), (what, f, rel_err(got[f], want[f]))
        else:
            assert rel_err(got[f], want[f]) <= 2e-6, (what, f, rel_err(got[f], want[f]))


@pytest.mark.parametrize("precision", [0, 1])
def test_visc_walk_tiles(precision):
    """The fused viscous pass walks `tile` rings per thread (1: the
    cell-parallel form) and, in an adiabatic disk, runs SubStep3 in the same
    walk from the tensor in registers.  Every tile, including ragged last
    tiles and tiles longer than the grid, gives the step of the separate
    stress, update and SubStep3 kernels (FP64: bit for bit; split: up to the
    separate kernels' FP32 tensors) on the adiabatic SN disk with viscous
    heating at 512 x 1536 (one step) and on the 32-ring fixture disk (5
    steps); the walk without SubStep3 stores the tensor of the separate
    stress kernel."""
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import KERNEL_STRESS, KERNELS, Disk
    name = "adi_512x1536"
    cfg = W.disk_config(name, precision)

    def big(tile, fused):
        d = Disk(cfg)
        W.start(d, name)
        d.set_fused_step(fused)
        d.set_visc_tile(tile)
        d.run_step()
        out = {f: d.download_field(f) for f in STEP_FIELDS}
        d.close()
        return out

    def tensor(tile):
        d = Disk(cfg)
        W.start(d, name)
        d.set_visc_tile(tile)
        if tile:
            d.time_kernel(KERNELS["visc_fused"], 1)
        else:
            d.time_kernel(KERNEL_STRESS, 1)
        out = {f: d.download_field(f) for f in TENSOR}
        d.close()
        return out

    want = big(0, False)
    want_t = tensor(0)
    for tile in (1, 3, 16, 64):
        print(f"viscous walk {name} tile={tile} precision={precision}", flush=True)
        _check(big(tile, True), want, precision, tile)
        _check(tensor(tile), want_t, precision, tile)
    meta, data = load_case("c2_small")
    c = dict(disk_config(meta), precision=precision)

    def small(tile, fused):
        d = Disk(c)
        d.set_fused_step(fused)
        d.set_visc_tile(tile)
        run_case(d, meta, data, nsteps=5)
        out = {f: d.download_field(f) for f in STEP_FIELDS}
        d.close()
        return out

    want = small(0, False)
    for tile in (3, 64):
        print(f"viscous walk c2_small tile={tile} precision={precision}", flush=True)
        _check(small(tile, True), want, precision, tile)


KICK_FIELDS = FIELDS + ["POTENTIAL"]


def _run_kick(meta, data, precision, kick_tile, nsteps, visc_tile=4):
    from fargocpt_amd.capi import Disk
    d = Disk(dict(disk_config(meta), precision=precision))
    d.set_visc_tile(visc_tile)
    d.set_kick_tile(kick_tile)
    run_case(d, meta, data, nsteps=nsteps)
    out = {f: d.download_field(f) for f in KICK_FIELDS}
    d.close()
    return out


@pytest.mark.parametrize("case", ["c1_small", "c2_small", "c5_small", "mix_small", "lf_iso", "lf_adi",
                                  "c1_iso"])
def test_kick_walk_fixtures(case):
    """The kick walk (sources, compression heating, artificial viscosity and
    its viscosity recalculation in one pass, then the viscous walk) against
    the same step with those as separate passes: every field bit for bit in FP64 storage after whole
    fixture runs (TW and SN, isothermal and adiabatic, Euler and LeapFrog),
    with ragged and single-tile walks."""
    meta, data = load_case(case)
    n = min(meta["nsteps"], 20)
    want = _run_kick(meta, data, 0, 1, n)
    for tile in (3, 64):
        print(f"kick walk case={case} tile={tile} steps={n}", flush=True)
        got = _run_kick(meta, data, 0, tile, n)
        for f in KICK_FIELDS:
            assert np.array_equal(got[f], want[f]), (tile, f, rel_err(got[f], want[f]))


SPLIT_VS_FP64 = 2e-5  # relative, as for the split-storage fixture runs


@pytest.mark.parametrize("name", ["adi_512x1536", "iso_2048x6144"])
def test_kick_walk_workloads(name):
    """Two steps of the benchmark disks with the size-based launch shapes
    against the same steps with the kick as separate passes: FP64 storage
    bit for bit; in split storage the walk keeps the source-term velocities
    in FP64 between its stages where the separate passes store them as FP32
    deviations, so both split runs are held to the split
    storage's tolerance against the FP64 run."""
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Disk

    def run(precision, tile):
        print(f"kick walk {name} precision={precision} tile={tile}", flush=True)
        d = Disk(W.disk_config(name, precision))
        W.start(d, name)
        d.set_kick_tile(tile)
        d.run_step()
        d.run_step()
        out = {f: d.download_field(f) for f in KICK_FIELDS}
        d.close()
        return out

    ref = run(0, 1)
    got = run(0, 0)
    for f in KICK_FIELDS:
        assert np.array_equal(got[f], ref[f]), (f, rel_err(got[f], ref[f]))
    for tile in (0, 1):
        got = run(1, tile)
        for f in ("SIGMA", "VRAD", "VAZI", "ENERGY", "POTENTIAL"):
            assert rel_err(got[f], ref[f]) <= SPLIT_VS_FP64, (tile, f, rel_err(got[f], ref[f]))


@pytest.mark.parametrize("integrator", [0, 1])
def test_fused_graph_matches_stepwise(integrator):
    """Graph replay of the fused step (both views captured) against fused
    steps launched one by one, on the adiabatic SN disk, Euler and LeapFrog."""
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Disk
    name = "adi_512x1536"
    print(f"{name}: fused graph vs run_step, integrator {integrator}", flush=True)
    cfg = dict(W.disk_config(name), integrator=integrator)
    a = Disk(cfg)
    b = Disk(cfg)
    for d in (a, b):
        W.start(d, name)
    a.run_steps(25, True)
    a.synchronize()
    for _ in range(25):
        b.run_step()
    for f in ("SIGMA", "VRAD", "VAZI", "ENERGY"):
        assert np.array_equal(a.download_field(f), b.download_field(f)), f
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_walks_keeps_state(precision):
    """fcpt_autotune_walks times whole steps per tile from the same state and
    puts the state and clock back: a tuned disk then steps exactly like an
    untuned one (the walks' tiles do not change the result in FP64; in split
    storage the same holds, the tiles only regroup the same per-cell work)."""
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Disk
    name = "adi_512x1536"
    print(f"autotune_walks {name} precision {precision}", flush=True)
    a, b = Disk(W.disk_config(name, precision)), Disk(W.disk_config(name, precision))
    for d in (a, b):
        W.start(d, name)
        d.run_steps(3, True)
        d.synchronize()
    t0 = a.get_time()
    tuned = a.autotune_walks(2)
    assert a.get_time() == t0
    assert tuned["kick_tile"] >= 1 and tuned["visc_tile"] >= 1
    for d in (a, b):
        d.run_steps(4, True)
        d.synchronize()
    assert a.get_time() == b.get_time()
    for f in ("SIGMA", "VRAD", "VAZI", "ENERGY", "SOUNDSPEED", "QPLUS"):
        assert np.array_equal(a.download_field(f), b.download_field(f)), f
    a.close()
    b.close()
