"""The C++ host interface (fcpt::sim::run / run_stepwise and the reference-named
functions of fargocpt_amd/csrc/fcpt_host.hpp) driven by tests/cpp/host_api_check
against the reference fixture of the 32 x 64 adiabatic disk (20 steps), within
1e-11 relative per field like the C-ABI runs."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from fixtures import disk_config, initial_condition, load_case, rel_err

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "host_api_check.bin")


@pytest.mark.parametrize("mode", ["stepwise", "device"])
@pytest.mark.parametrize("case", ["c2_small", "mix_small"])
def test_cpp_host_interface(case, mode):
    from fargocpt_amd.capi import desc_from_dict
    assert os.path.exists(EXE), "build tests/cpp/host_api_check with __graft_entry__.build()"
    meta, data = load_case(case)
    cfg = disk_config(meta)
    desc, radii = desc_from_dict(cfg)
    ic = initial_condition(meta)
    n = meta["nsteps"]
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "desc.bin"), "wb").write(bytes(desc))
        radii.tofile(os.path.join(d, "radii.bin"))
        for f in ("sigma", "vrad", "vazi", "energy"):
            ic[f].tofile(os.path.join(d, f + ".bin"))
        rec = []
        for s in range(n):
            rec.append(data["planets"][s].ravel())
            rec.append(data["steps"][s][6:8])
        np.concatenate(rec).tofile(os.path.join(d, "bodies.bin"))
        print(f"case={case} mode={mode}", flush=True)
        subprocess.run([EXE, d, str(n), mode], check=True, timeout=120)
        t, last = np.fromfile(os.path.join(d, "out_time.bin"))
        assert abs(t - data["steps"][n - 1][8]) <= 1e-11 * t
        for f, shape in (("sigma", (meta["Nr"], meta["Nphi"])),
                         ("vrad", (meta["Nr"] + 1, meta["Nphi"])),
                         ("vazi", (meta["Nr"], meta["Nphi"])),
                         ("energy", (meta["Nr"], meta["Nphi"]))):
            got = np.fromfile(os.path.join(d, f"out_{f}.bin")).reshape(shape)
            assert rel_err(got, data[f"state_{n}_{f}"]) < 1e-11, f
