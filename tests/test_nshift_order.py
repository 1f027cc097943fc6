"""The FARGO shift's dependence on the ring sum's order, on the CPU: the
constructed rings of nshift_cases.py really straddle a half-integer Ntilde
between the two orders, and the oracle (the reference's left-to-right sum)
takes the sequential side."""
import numpy as np
import pytest

from nshift_cases import STEP_PLANETS, build_case, quiet_config, seq_sum, tree_sum
from oracle import OracleDisk


@pytest.mark.parametrize("kind", ["cancel", "ulp"])
def test_constructed_ring_straddles(kind, oracle_lib):
    cfg, st, ring, dt, n_seq, n_tree = build_case(kind)
    assert n_seq != n_tree
    d = OracleDisk(cfg)
    d.upload_state(st["sigma"], st["vrad"], st["vazi"], None)
    d.transport(dt)
    assert d.nshift()[ring] == n_seq
    d.close()


def test_tree_sum_emulation_matches_python_for_exact_values():
    # integer-valued terms: every order gives the exact sum
    v = np.arange(1, 1001, dtype=np.float64)
    assert tree_sum(v) == seq_sum(v) == 500500.0


@pytest.mark.parametrize("kind", ["cancel", "ulp"])
def test_quiet_step_keeps_constructed_ring(kind, oracle_lib):
    """With no viscosity, no artificial viscosity, no damping and no bodies,
    the stages before the transport leave v_phi exactly as uploaded, so a
    whole Euler step meets the constructed ring unchanged."""
    cfg, st, ring, dt, n_seq, n_tree = build_case(kind)
    cfg = quiet_config(cfg)
    d = OracleDisk(cfg)
    d.upload_state(st["sigma"], st["vrad"], st["vazi"], None)
    d.set_step_params(STEP_PLANETS)
    d.nbody_potential()
    d.update_with_sourceterms(dt)
    d.artificial_viscosity(dt)
    d.recalculate_viscosity()
    d.viscous_stress_tensor()
    d.update_velocities_with_viscosity(dt)
    d.apply_boundary_condition(dt, False)
    assert np.array_equal(d.download_field("VAZI")[1:-1], st["vazi"][1:-1])
    d.upload_state(st["sigma"], st["vrad"], st["vazi"], None)
    d.step_euler(dt)
    assert d.nshift()[ring] == n_seq
    d.close()
