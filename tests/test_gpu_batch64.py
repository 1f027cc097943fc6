"""The 64-disk batched sweep at its full size (Nr uniform in [128, 1024],
Nphi = 3 Nr, q, alpha and h drawn per disk): every member batched must evolve
bit for bit like the same disk stepped alone with fcpt_run_steps, in FP64 and
in split storage; and the three smallest members, whose alpha / h / q draws
lie outside every reference fixture, must follow the oracle step by step in
FP64: identical FARGO shifts at every step, fields within 1e-11 relative."""
import numpy as np
import pytest

from fixtures import rel_err

pytestmark = pytest.mark.gpu

STEPS = 20
FIELDS = ("SIGMA", "VRAD", "VAZI")


@pytest.mark.parametrize("prec", [0, 1])
def test_batch64_full_size_matches_single_runs(prec):
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Batch, Disk
    members = W.batch_members()
    print(f"batch64 precision {prec}: {len(members)} disks, "
          f"{sum(m['nr'] * m['nphi'] for m in members)} cells", flush=True)
    alone = []
    for m in members:
        d = Disk(W.member_config(m, prec))
        W.member_start(d, m)
        d.run_steps(STEPS, True)
        d.synchronize()
        alone.append(({f: d.download_field(f) for f in FIELDS}, d.get_time(), d.nshift()))
        d.close()
    print("single runs done", flush=True)
    disks = []
    for m in members:
        d = Disk(W.member_config(m, prec))
        W.member_start(d, m)
        disks.append(d)
    b = Batch(disks)
    b.run_steps(STEPS, True)
    b.synchronize()
    print("batch done", flush=True)
    for k, (d, (fields, clock, nsh)) in enumerate(zip(disks, alone)):
        for f in FIELDS:
            assert np.array_equal(d.download_field(f), fields[f]), (k, members[k]["nr"], f)
        assert d.get_time() == clock, k
        assert np.array_equal(d.nshift(), nsh), k
    b.close()
    for d in disks:
        d.close()


def test_batch64_smallest_members_follow_oracle():
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Disk
    from oracle import OracleDisk
    members = sorted(W.batch_members(), key=lambda m: m["nr"])[:3]
    for m in members:
        print(f"member nr={m['nr']} q={m['q']:.3e} alpha={m['alpha']:.3e} h={m['h']:.4f}",
              flush=True)
        gpu = Disk(W.member_config(m, 0))
        cpu = OracleDisk(W.member_config(m, 0))
        for d in (gpu, cpu):
            W.member_start(d, m)
        for s in range(STEPS):
            gpu.run_steps(1, True)
            gpu.synchronize()
            cpu.run_step()
            assert np.array_equal(gpu.nshift(), cpu.nshift()), (m["nr"], s)
            assert abs(gpu.get_time()[0] - cpu.get_time()[0]) <= 1e-12 * cpu.get_time()[0]
        for f in FIELDS:
            err = rel_err(gpu.download_field(f), cpu.download_field(f))
            assert err <= 1e-11, (m["nr"], f, err)
        gpu.close()
        cpu.close()
