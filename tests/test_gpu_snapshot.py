"""Snapshot files through the C-ABI against the files the reference's own
writers produced for the same state (t_polargrid::write2D, output::write_misc,
the used_rad.dat and dimensions.dat writers), kept byte for byte in the
fixtures, and restarts from them.

The state is uploaded from the fixture (an exact FP64 round trip), so the
grids must come out byte-identical; misc.bin takes the run counters given to
the writer and the context's clock.  Derived grids written from the device
(Temperature, pressure) are compared by value.
"""
import struct

import numpy as np
import pytest

from fixtures import disk_config, load_case, rel_err, step_inputs

pytestmark = pytest.mark.gpu

GRIDS = ("Sigma.dat", "vrad.dat", "vazi.dat", "energy.dat", "Qplus.dat", "Qminus.dat")
STATE = ("sigma", "vrad", "vazi", "energy", "qplus", "qminus")


def _misc(data):
    snap, nmon, t, omega, angle, last_dt, n_iter = struct.unpack("<IIddddQ", bytes(data["snap_misc.bin"]))
    return dict(snapshot=snap, n_monitor=nmon, time=t, omega_frame=omega, frame_angle=angle,
                last_dt=last_dt, n_iter=n_iter)


def _load_final(d, meta, data):
    n = meta["nsteps"]
    st = {f: data[f"state_{n}_{f}"] for f in STATE}
    d.upload_reference(*(data["init_" + f] for f in ("sigma0", "vrad0", "vazi0", "energy0")))
    d.upload_state(st["sigma"], st["vrad"], st["vazi"], st["energy"])
    d.upload_field("QPLUS", st["qplus"])
    d.upload_field("QMINUS", st["qminus"])
    m = _misc(data)
    d.set_time(m["time"], m["last_dt"], m["n_monitor"])
    return st, m


@pytest.mark.parametrize("case", ["c1_small", "c2_small"])
def test_write_snapshot_matches_reference_files(case, tmp_path):
    from fargocpt_amd.capi import Disk
    meta, data = load_case(case)
    print(f"snapshot write case={case}", flush=True)
    d = Disk(disk_config(meta))
    _, m = _load_final(d, meta, data)
    d.write_snapshot(tmp_path, snapshot=m["snapshot"], omega_frame=m["omega_frame"],
                     frame_angle=m["frame_angle"], n_iter=m["n_iter"],
                     fields=("SIGMA", "VRAD", "VAZI", "ENERGY", "QPLUS", "QMINUS",
                             "TEMPERATURE", "PRESSURE"))
    for name in GRIDS + ("misc.bin",):
        assert (tmp_path / name).read_bytes() == bytes(data["snap_" + name]), name
    for name in ("Temperature.dat", "pressure.dat"):
        got = np.fromfile(tmp_path / name)
        want = np.frombuffer(bytes(data["snap_" + name]), dtype=np.float64)
        assert got.shape == want.shape and rel_err(got, want) <= 1e-14, name
    d.write_grid_files(tmp_path, "Logarithmic")
    for name in ("used_rad.dat", "dimensions.dat"):
        assert (tmp_path / name).read_bytes() == bytes(data["grid_" + name]), name
    d.close()


@pytest.mark.parametrize("case", ["c1_small", "c2_small"])
def test_restart_from_reference_snapshot(case, tmp_path):
    """Reading the reference's snapshot gives the state, the clock and the
    counters it holds, and the run continues exactly as from uploading the
    same state."""
    from fargocpt_amd.capi import Disk
    meta, data = load_case(case)
    print(f"snapshot restart case={case}", flush=True)
    snap = tmp_path / "snap"
    ref = tmp_path / "reference"
    snap.mkdir()
    ref.mkdir()
    for name in GRIDS + ("misc.bin",):
        (snap / name).write_bytes(bytes(data["snap_" + name]))
    for name, key in zip(("Sigma.dat", "vrad.dat", "vazi.dat", "energy.dat"),
                         ("sigma0", "vrad0", "vazi0", "energy0")):
        data["init_" + key].tofile(ref / name)
    cfg = disk_config(meta)
    a = Disk(cfg)
    info = a.read_snapshot(snap)
    a.read_snapshot_reference(ref)
    assert info == _misc(data)
    n = meta["nsteps"]
    for f in STATE if meta["adiabatic"] else STATE[:3]:
        assert np.array_equal(a.download_field(f.upper()), data[f"state_{n}_{f}"]), f
    for f, key in (("SIGMA0", "sigma0"), ("VRAD0", "vrad0"), ("VAZI0", "vazi0")):
        assert np.array_equal(a.download_field(f), data["init_" + key]), f
    t, last_dt, nmon = a.get_time()
    assert (t, last_dt, nmon) == (info["time"], info["last_dt"], info["n_monitor"])
    b = Disk(cfg)
    _load_final(b, meta, data)
    planets, ind = step_inputs(data, n)
    for x in (a, b):
        x.set_step_params(planets, ind)
        for _ in range(3):
            x.run_step()
    for f in ("SIGMA", "VRAD", "VAZI", "ENERGY"):
        assert np.array_equal(a.download_field(f), b.download_field(f)), f
    a.close()
    b.close()


def test_snapshot_rejects_wrong_size(tmp_path):
    from fargocpt_amd.capi import Disk
    meta, data = load_case("c1_small")
    d = Disk(disk_config(meta))
    for name in GRIDS + ("misc.bin",):
        (tmp_path / name).write_bytes(bytes(data["snap_" + name]))
    (tmp_path / "vazi.dat").write_bytes(bytes(data["snap_vazi.dat"])[:-8])
    with pytest.raises(RuntimeError, match="bytes"):
        d.read_snapshot(tmp_path)
    d.close()
