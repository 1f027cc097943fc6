"""The FARGO integer shifts on the GPU follow the reference's left-to-right
ring sum (TransportEuler.cpp:178-183, 215-220) by construction: rings whose
tree sum and left-to-right sum give different shifts (nshift_cases.py) get
the left-to-right one, through the per-function transport (k_ring_mean) and
through the fused Euler step (the ghost-ring + ring-mean kernel), in FP64
storage.  Every ring's shift must equal the oracle's."""
import numpy as np
import pytest

from nshift_cases import STEP_PLANETS, build_case, quiet_config
from oracle import OracleDisk

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["cancel", "ulp"])
def test_transport_shift_follows_reference_order(kind):
    from fargocpt_amd.capi import Disk
    cfg, st, ring, dt, n_seq, n_tree = build_case(kind)
    print(f"kind={kind} ring={ring} dt={dt!r} seq={n_seq} tree={n_tree}", flush=True)
    assert n_seq != n_tree
    shifts = []
    for d in (Disk(cfg), OracleDisk(cfg)):
        d.upload_state(st["sigma"], st["vrad"], st["vazi"], None)
        d.transport(dt)
        shifts.append(d.nshift())
        d.close()
    gpu, cpu = shifts
    assert gpu[ring] == n_seq
    assert np.array_equal(gpu, cpu)


@pytest.mark.parametrize("kind", ["cancel", "ulp"])
def test_fused_step_shift_follows_reference_order(kind):
    from fargocpt_amd.capi import Disk
    cfg, st, ring, dt, n_seq, n_tree = build_case(kind)
    cfg = quiet_config(cfg)
    print(f"kind={kind} ring={ring} dt={dt!r} seq={n_seq} tree={n_tree} (fused step)", flush=True)
    shifts = []
    for d in (Disk(cfg), OracleDisk(cfg)):
        d.upload_state(st["sigma"], st["vrad"], st["vazi"], None)
        d.set_step_params(STEP_PLANETS)
        d.step_euler(dt)
        shifts.append(d.nshift())
        d.close()
    gpu, cpu = shifts
    assert gpu[ring] == n_seq
    assert np.array_equal(gpu, cpu)
