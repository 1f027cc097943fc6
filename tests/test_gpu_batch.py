"""Batched disks (fcpt_batch_*): several disks of different sizes and physics
advanced by one launch per step stage must evolve bit for bit like the same
disks stepped alone with fcpt_run_steps (the same blocks do the same
arithmetic, with the walks where the disk's own launches use them), in FP64
and in split storage."""
import numpy as np
import pytest

from fixtures import disk_config, load_case, start_run, step_inputs

pytestmark = pytest.mark.gpu

CASES = ["c1_small", "c2_small", "c5_small", "mix_small", "none_small", "c1_iso"]
FIELDS = ["SIGMA", "VRAD", "VAZI", "ENERGY", "PRESSURE", "SOUNDSPEED", "VISCOSITY", "QPLUS"]


def _fixture_disks(prec):
    from fargocpt_amd.capi import Disk
    disks = []
    for case in CASES:
        meta, data = load_case(case)
        cfg = disk_config(meta)
        cfg["precision"] = prec
        d = Disk(cfg)
        start_run(d, meta)
        planets, ind = step_inputs(data, 1)
        d.set_step_params(planets, ind)
        disks.append((d, meta))
    return disks


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_matches_single_disks(prec):
    from fargocpt_amd.capi import Batch
    print(f"batch of {len(CASES)} fixture disks, precision {prec}", flush=True)
    single = _fixture_disks(prec)
    batched = _fixture_disks(prec)
    for d, _ in single:
        d.run_steps(15, True)
        d.synchronize()
    b = Batch([d for d, _ in batched])
    b.run_steps(5, True)
    b.run_steps(10, False)
    b.synchronize()
    for (s, meta), (d, _) in zip(single, batched):
        for f in FIELDS:
            if f in ("ENERGY", "QPLUS") and not meta["adiabatic"]:
                continue
            assert np.array_equal(s.download_field(f), d.download_field(f)), (meta["Nr"], f)
        assert s.get_time() == d.get_time()
    b.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_walks_match_single_disks(prec):
    """The fixture disks with the walks forced on (3 rings per thread for
    the kick and viscous walks, ragged last tiles): the batch runs the walks'
    batched forms where each disk's own launch runs the walks, and every disk
    evolves bit for bit like it does alone."""
    from fargocpt_amd.capi import Batch
    print(f"batch of {len(CASES)} fixture disks with walks, precision {prec}", flush=True)
    single = _fixture_disks(prec)
    batched = _fixture_disks(prec)
    for d, _ in single + batched:
        d.set_kick_tile(3)
        d.set_visc_tile(3)
    for d, _ in single:
        d.run_steps(12, True)
        d.synchronize()
    b = Batch([d for d, _ in batched])
    b.run_steps(12, True)
    b.synchronize()
    for (s, meta), (d, _) in zip(single, batched):
        for f in FIELDS:
            if f in ("ENERGY", "QPLUS") and not meta["adiabatic"]:
                continue
            assert np.array_equal(s.download_field(f), d.download_field(f)), (meta["Nr"], f)
        assert s.get_time() == d.get_time()
    b.close()


def test_batch_members_uneven_sizes():
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Batch, Disk
    members = sorted(W.batch_members(), key=lambda m: m["nr"])[:6]
    print("batch of the six smallest batch64 members:", [m["nr"] for m in members], flush=True)
    single, batched = [], []
    for m in members:
        for lst in (single, batched):
            d = Disk(W.member_config(m, 1))
            W.member_start(d, m)
            lst.append(d)
    for d in single:
        d.run_steps(20, True)
        d.synchronize()
    b = Batch(batched)
    b.run_steps(20, True)
    b.synchronize()
    for s, d in zip(single, batched):
        for f in ("SIGMA", "VRAD", "VAZI"):
            assert np.array_equal(s.download_field(f), d.download_field(f)), f
        m0, _ = s.totals(1e300)
        assert np.isfinite(m0) and m0 > 0
    b.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_autotune_keeps_results(prec):
    """The batch's radial-sweep autotuner changes launch shapes only: after
    it, every disk still evolves bit for bit as alone."""
    from fargocpt_amd.capi import Batch
    print(f"batch autotune, precision {prec}", flush=True)
    single = _fixture_disks(prec)
    batched = _fixture_disks(prec)
    for d, _ in single:
        d.run_steps(8, True)
        d.synchronize()
    b = Batch([d for d, _ in batched])
    b.run_steps(3, True)
    b.synchronize()
    shapes = b.autotune(2)
    assert shapes["radial_tile"] in (4, 8, 16, 32) and shapes["radial_waves"] in (2, 3)
    b.run_steps(5, True)
    b.synchronize()
    for (s, meta), (d, _) in zip(single, batched):
        for f in FIELDS:
            if f in ("ENERGY", "QPLUS") and not meta["adiabatic"]:
                continue
            assert np.array_equal(s.download_field(f), d.download_field(f)), (meta["Nr"], f)
        assert s.get_time() == d.get_time()
    b.close()
