"""fcpt_totals (quantities::gas_total_mass / gas_angular_momentum,
quantities.cpp:51-78, 242-279) on the GPU against the reference's own totals
of the same state, recorded by the fixture driver with one OpenMP thread.

FP64 storage: within the bound between a compensated and a left-to-right
sum, (n + 2) u sum |term| (u = 2^-53).  Split storage: the uploaded state is
rounded to FP32 deviations about the ring means, so the GPU's totals are
checked against the compensated totals of the state the disk holds (the
download, background + deviation in FP64) within the same bound, and against
the reference within the storage rounding: each deviation is rounded to
2^-24 of itself and a ring's deviations sum to at most twice its values'
magnitudes, so 2^-22 sum |term| covers both factors of a term."""
import numpy as np
import pytest

from fixtures import disk_config, load_case, totals_bound, totals_magnitudes

pytestmark = pytest.mark.gpu

CASES = [("c1_small", "20"), ("c2_small", "5"), ("c5_small", "1"), ("c5_small", "20"),
         ("mix_small", "20"), ("c1_iso", "100"), ("lf_adi", "20")]


def _compensated(sigma, vazi, surf, rmed, om, radius, nr):
    """Neumaier sums of the two totals in numpy scalars (the exact sum to 2u)."""
    m = mc = l = lc = 0.0
    for i in range(1, nr - 1):
        if not rmed[i] <= radius:
            continue
        s = sigma[i]
        dm = surf[i] * s
        dl = surf[i] * 0.5 * (s + np.roll(s, 1)) * rmed[i] * (vazi[i] + om * rmed[i])
        for x in dm.tolist():
            t = m + x
            mc += (m - t) + x if abs(m) >= abs(x) else (x - t) + m
            m = t
        for x in dl.tolist():
            t = l + x
            lc += (l - t) + x if abs(l) >= abs(x) else (x - t) + l
            l = t
    return m + mc, l + lc


@pytest.mark.parametrize("case,tag", CASES)
@pytest.mark.parametrize("precision", [0, 1])
def test_totals_vs_reference(case, tag, precision):
    from fargocpt_amd.capi import Disk
    meta, data = load_case(case)
    print(f"case={case} checkpoint={tag} precision={precision}", flush=True)
    cfg = disk_config(meta)
    cfg["precision"] = precision
    d = Disk(cfg)
    nr = meta["Nr"]
    surf, rmed, om = data["geom_Surf"], data["geom_Rmed"], meta["OmegaFrame"]
    sig, vaz = data[f"state_{tag}_sigma"], data[f"state_{tag}_vazi"]
    energy = data[f"state_{tag}_energy"] if meta["adiabatic"] else None
    d.upload_state(sig, data[f"state_{tag}_vrad"], vaz, energy)
    ref = data["totals"][int(tag) - 1]
    for radius, m_ref, l_ref in ((1e300, ref[0], ref[1]), (ref[4], ref[2], ref[3])):
        m, l = d.totals(radius)
        am, al, n = totals_magnitudes(sig, vaz, surf, rmed, om, radius, nr)
        if precision == 0:
            assert abs(m - m_ref) <= totals_bound(n, am), (radius, m, m_ref)
            assert abs(l - l_ref) <= totals_bound(n, al), (radius, l, l_ref)
        else:
            s_held, v_held = d.download_field("SIGMA"), d.download_field("VAZI")
            m_held, l_held = _compensated(s_held, v_held, surf, rmed, om, radius, nr)
            am_h, al_h, _ = totals_magnitudes(s_held, v_held, surf, rmed, om, radius, nr)
            assert abs(m - m_held) <= totals_bound(n, am_h), (radius, m, m_held)
            assert abs(l - l_held) <= totals_bound(n, al_h), (radius, l, l_held)
            assert abs(m - m_ref) <= 2.0 ** -22 * am, (radius, m, m_ref)
            assert abs(l - l_ref) <= 2.0 ** -22 * al, (radius, l, l_ref)
    d.close()
