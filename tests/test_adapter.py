"""The reference-side adapter (tests/cpp/gpu_hydro.cpp) compiles against the
reference's own headers and links with the reference's objects and
libfcpt_hip.so: a renamed parameters:: / boundary_conditions:: / refframe::
symbol, a changed planet accessor or a changed fcpt_* signature fails here.
Needs the reference sources (/root/reference/src) and the reference objects
built by oracle/ref.mk; skipped where they are absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
REF_OUT = os.path.join(ROOT, "oracle", "_ref")
LIBS = ("libfargo_ref.a", "librebound.a", "libyamlcpp.a", "libunits.a")
FLAGS = ["-O1", "-std=c++17", "-DNDEBUG", "-DOMPI_SKIP_MPICXX", "-DDISABLE_FFTW", "-DDISABLE_GSL",
         f"-I{REF_SRC}", "-I/opt/conda/include", f"-I{os.path.join(ROOT, 'include')}", "-fopenmp",
         "-no-pie", "-Wall", "-Wno-unused-function"]


def _have_reference():
    return (os.path.isdir(REF_SRC) and all(os.path.exists(os.path.join(REF_OUT, x)) for x in LIBS)
            and os.path.exists(os.path.join(ROOT, "fargocpt_amd", "libfcpt_hip.so")))


@pytest.mark.skipif(not _have_reference(), reason="reference sources or objects absent")
def test_adapter_compiles_and_links(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "gpu_hydro.cpp")
    main = os.path.join(ROOT, "tests", "cpp", "gpu_hydro_link.cpp")
    obj, mobj, exe = (str(tmp_path / n) for n in ("gpu_hydro.o", "main.o", "adapter_check"))
    r = subprocess.run(["g++", *FLAGS, "-c", src, "-o", obj], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "warning" not in r.stderr, r.stderr[-4000:]
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-c", main, "-o", mobj], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    lib = os.path.join(ROOT, "fargocpt_amd")
    mpilink = os.path.join(REF_OUT, "mpilib")
    r = subprocess.run(["g++", "-no-pie", "-o", exe, mobj, obj,
                        *[os.path.join(REF_OUT, x) for x in LIBS],
                        f"-L{lib}", "-lfcpt_hip", f"-Wl,-rpath-link,{lib}",
                        "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
                        f"-L{mpilink}", f"-Wl,-rpath-link,{mpilink}", "-lmpi", "-lm", "-lrt",
                        "-fopenmp"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    nm = subprocess.run(["nm", "-u", exe], capture_output=True, text=True).stdout
    used = {ln.split()[-1] for ln in nm.splitlines() if ln.split() and ln.split()[-1].startswith("fcpt_")}
    # every C-ABI function the adapter calls is one the header declares
    header = open(os.path.join(ROOT, "include", "fcpt.h")).read()
    assert used and all(f"{name}(" in header for name in used), sorted(used)
