"""Error reporting of the step entry points.

A state with one radial velocity of 1e300 gives that cell a CFL limit of 0
(its squared rate overflows to +inf; an infinite velocity would make the
cell's limit NaN, which never wins the minimum), so the reduction's result
is not positive and fcpt_run_step reports FCPT_ERR_NONFINITE.  The condition is forced through the inputs: every
kernel runs to its end on finite, infinite or NaN values (the step that
follows the report runs with dt = 0, its FARGO shifts 0), none with a
data-dependent index.  The report is one-shot: once a good state and clock
are uploaded, the next steps run clean.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_nonpositive_dt_reported_then_cleared():
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Disk
    from fargocpt_amd.ic import initial_state
    name = "iso_128x384"
    print(f"{name}: one v_r of 1e300, then the good state", flush=True)
    d = Disk(W.disk_config(name, 0))
    ic = initial_state(W.setup(name))
    d.upload_reference(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    bad_vr = ic["vrad"].copy()
    bad_vr[40, 17] = 1e300
    d.upload_state(ic["sigma"], bad_vr, ic["vazi"], ic["energy"])
    d.set_time(0.0, 0.1, 0)
    d.set_step_params([(0.0, 0.0, 1.0, 0.0)], (0.0, 0.0))
    with pytest.raises(RuntimeError, match=r"\(-4\)"):
        d.run_step()
    # the good state and clock: no report left over
    d.upload_state(ic["sigma"], ic["vrad"], ic["vazi"], ic["energy"])
    d.set_time(0.0, 0.1, 0)
    for _ in range(2):
        dt = d.run_step()
        assert np.isfinite(dt) and dt > 0.0
    for f in ("SIGMA", "VRAD", "VAZI"):
        assert np.all(np.isfinite(d.download_field(f))), f
    d.close()
