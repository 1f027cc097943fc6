"""The C-ABI library: it loads without a GPU and exports every function that
include/fcpt.h declares; argument checks happen before any device call."""
import ctypes as C
import os

import numpy as np

from fargocpt_amd import capi


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    names = capi.exported_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layout_matches_header():
    """Field order of the ctypes mirror follows include/fcpt.h."""
    import re
    hdr = open(os.path.join(os.path.dirname(capi.HERE), "include", "fcpt.h")).read()
    body = hdr[hdr.index("typedef struct fcpt_disk_desc {"):hdr.index("} fcpt_disk_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names, types = [], []
    for decl in re.findall(r"(int32_t|const double \*|double)\s*([^;]+);", body):
        for n in decl[1].split(","):
            names.append(n.strip().lstrip("*"))
            types.append(decl[0])
    assert names == [n for n, _ in capi.DiskDesc._fields_]
    ct = {"int32_t": C.c_int32, "double": C.c_double}
    for (n, t), ht in zip(capi.DiskDesc._fields_, types):
        if ht in ct:
            assert t is ct[ht], n


def test_create_rejects_bad_sizes_before_touching_the_device():
    lib = capi.load_library()
    radii = np.linspace(0.4, 2.5, 7)
    d = capi.DiskDesc()
    d.nr, d.nphi = 5, 16  # fewer rings than the minimum
    d.radii = radii.ctypes.data_as(C.POINTER(C.c_double))
    h = C.c_void_p()
    assert lib.fcpt_create(C.byref(d), C.byref(h)) == -1
    d.nr = 20000  # beyond the reference's MAX1D
    assert lib.fcpt_create(C.byref(d), C.byref(h)) == -1
    assert lib.fcpt_create(None, C.byref(h)) == -1


def test_batch_members_are_reproducible():
    from fargocpt_amd import workloads as W
    a, b = W.batch_members(), W.batch_members()
    assert a == b and len(a) == 64
    for m in a:
        assert 128 <= m["nr"] <= 1024 and m["nphi"] == 3 * m["nr"]
        assert 1e-5 <= m["q"] <= 1e-3 and 1e-4 <= m["alpha"] <= 1e-2 and 0.03 <= m["h"] <= 0.07
    cfg = W.member_config(a[0], 1)
    assert cfg["nr"] == a[0]["nr"] and cfg["precision"] == 1 and len(cfg["radii"]) == a[0]["nr"] + 2


def test_totals_noalloc_rejects_bad_sizes_before_touching_the_device():
    lib = capi.load_library()
    geo = np.ones(20000)
    m, l = C.c_double(), C.c_double()
    dp = geo.ctypes.data_as(C.POINTER(C.c_double))
    fake = C.c_void_p(0x1000)  # never dereferenced: the size check comes first
    rc = lib.fcpt_totals_noalloc(20000, 64, dp, dp, fake, fake, 1.0, 1e300, None,
                                 C.byref(m), C.byref(l))
    assert rc == -1
    assert b"exceeds" in lib.fcpt_noalloc_last_error()
    rc = lib.fcpt_totals_noalloc(128, 64, dp, dp, None, fake, 1.0, 1e300, None, C.byref(m), C.byref(l))
    assert rc == -1
