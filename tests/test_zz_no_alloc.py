"""fcpt_totals_noalloc, the totals entry point that allocates nothing: over
caller-owned device copies of a disk's Sigma and v_phi it must return what
fcpt_totals returns for that disk, bit for bit (the same kernels and
summation order), on two calls in a row through the one static scratch array,
the second with a smaller disk.  The device arrays are the caller's: this test
allocates them with the HIP runtime the library itself is linked against."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime already loaded by
    libfcpt_hip.so (resolved by its soname, so the same instance)."""

    def __init__(self):
        from fargocpt_amd import capi
        capi.load_library()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # H2D
        return p.value

    def free(self, p: int):
        assert self.rt.hipFree(C.c_void_p(p)) == 0


def _disk(name, steps):
    from fargocpt_amd import workloads as W
    from fargocpt_amd.capi import Disk
    d = Disk(W.disk_config(name, 0))
    W.start(d, name)
    d.run_steps(steps, True)
    d.synchronize()
    return d, W.disk_config(name)["omega_frame"]


def test_totals_noalloc_matches_fcpt_totals_twice():
    from fargocpt_amd.capi import totals_noalloc
    hip = _Hip()
    for name, steps in (("adi_512x1536", 3), ("iso_128x384", 5)):  # the second is smaller
        print(f"no-alloc totals on {name} after {steps} steps", flush=True)
        d, omega = _disk(name, steps)
        rmed, surf = d.ring_geometry()
        sig = hip.upload(d.download_field("SIGMA"))
        vaz = hip.upload(d.download_field("VAZI"))
        try:
            for radius in (1e300, 1.5):
                want = d.totals(radius)
                got = totals_noalloc(d.nr, d.nphi, rmed, surf, sig, vaz, omega, radius)
                assert got == want, (name, radius, got, want)
        finally:
            hip.free(sig)
            hip.free(vaz)
            d.close()
