"""The drop-in boundary in the reference's own C++ types.

scdat_amd/facade/PICCycle_device.cpp defines the members of
SCDAT::PICcore::PICCycle, the class declared in the reference's unchanged
Tools/PICcore/include/PICCycle.h, over the C ABI; a reference build links it
(and libscdat_amd.so) instead of Tools/PICcore/src/PICCycle.cpp.  Here the
file and two of the reference's callers of PICCycle -- the PIC benchmark
driver (Main/main_PICcore.cpp) and the surface-charging toolkit's live-PIC
sampler (PicMccSurfaceCurrentSampler.cpp) -- are compiled, unchanged,
against the reference's headers (oracle/ref/Makefile, target dropin), and
the symbol tables show that the replacement defines every PICCycle member
those callers need and nothing else a reference link could clash with.
Compile only: the reference's ParticleManager needs <format>, which this
image's standard library lacks, so no reference program is linked or run.
scdat_amd/facade/PoissonSolver_device.cpp, the device's
SCDAT::FieldSolver::PoissonSolver, replaces Tools/FieldSolver/src/
PoissonSolver.cpp the same way: it defines what the reference's file defines
and nothing else, and, linked with the reference's mesh sources, it refuses
to solve where no GPU is visible (no host fallback); its GPU parity is in
tests/test_gpu_poisson_dropin.py.  Skipped where the reference sources are
absent."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
DROP = os.path.join(ROOT, "oracle", "_ref", "dropin")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "Tools", "PICcore")),
                                reason="reference sources not present")


def _symbols(obj, flags):
    out = subprocess.run(["nm", "-C"] + flags + [obj], capture_output=True, text=True, check=True).stdout
    rows = []
    for line in out.splitlines():
        m = re.match(r"^\s*(?:[0-9a-f]+\s+)?([A-Za-z])\s+(.*)$", line)
        if m:
            rows.append((m.group(1), m.group(2)))
    return rows


@pytest.fixture(scope="module")
def objects():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref"), "all", "dropin"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    objs = {k: os.path.join(DROP, k + ".o")
            for k in ("PICCycle_device", "main_PICcore", "PicMccSurfaceCurrentSampler", "PoissonSolver_device")}
    objs["PoissonSolver_reference"] = os.path.join(ROOT, "oracle", "_ref", "obj", "FieldSolver", "src", "PoissonSolver.o")
    return objs


def test_replacement_defines_every_member_the_reference_callers_use(objects):
    needed = set()
    for caller in ("main_PICcore", "PicMccSurfaceCurrentSampler"):
        for kind, name in _symbols(objects[caller], ["--undefined-only"]):
            if name.startswith("SCDAT::PICcore::PICCycle::"):
                needed.add(name)
    assert len(needed) >= 8, needed  # constructor, setMesh, setParticleManager, ... executeTimeStep
    defined = {name for kind, name in _symbols(objects["PICCycle_device"], ["--defined-only"]) if kind == "T"}
    missing = sorted(needed - defined)
    assert not missing, missing


def test_replacement_defines_only_picycle_members(objects):
    # every strong global definition is a member of PICCycle: a link with the
    # reference's other objects cannot clash (helpers are file-local, template
    # instances weak)
    strong = [name for kind, name in _symbols(objects["PICCycle_device"], ["--defined-only", "--extern-only"])
              if kind in "TDBR"]
    others = [n for n in strong if not n.startswith("SCDAT::PICcore::PICCycle::")]
    assert not others, others
    for member in ("initialize()", "executeTimeStep()", "executeSimulation(double)", "updateElectricField()",
                   "clearBoundaryConditions()", "findParticleElement(SCDAT::Particle::Particle*) const"):
        assert any(n == "SCDAT::PICcore::PICCycle::" + member for n in strong), member


POISSON = ("SCDAT::FieldSolver::PoissonSolver::", "SCDAT::FieldSolver::BoundaryCondition::")


def test_poisson_replacement_defines_what_the_reference_file_defines(objects):
    ref = {name for kind, name in _symbols(objects["PoissonSolver_reference"], ["--defined-only", "--extern-only"])
           if kind == "T" and name.startswith(POISSON)}
    assert len(ref) >= 30, ref
    mine = {name for kind, name in _symbols(objects["PoissonSolver_device"], ["--defined-only", "--extern-only"])
            if kind == "T"}
    # the host least-squares helper is private and unused here: the device
    # computes the nodal gradient
    missing = sorted(n for n in ref - mine if "computeLeastSquaresGradient" not in n)
    assert not missing, missing
    strong = [name for kind, name in _symbols(objects["PoissonSolver_device"], ["--defined-only", "--extern-only"])
              if kind in "TDBR"]
    others = [n for n in strong if not n.startswith(POISSON)]
    assert not others, others


def test_poisson_replacement_refuses_without_a_device(objects):
    import ctypes as C

    import numpy as np

    import scdat_amd as S
    try:
        S.device_info()
        pytest.skip("a GPU is visible: the device solve is checked in tests/test_gpu_poisson_dropin.py")
    except S.ScdatError:
        pass
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libscdat_poisson_dropin.so"))
    cap = 1000
    n = C.c_int(0)
    z, phi, E, tot = np.zeros(cap), np.zeros(cap), np.zeros(3 * cap), np.zeros(3)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    rc = L.ref_poisson_plate(cap, C.byref(n), P(z), P(phi), P(E), P(tot))
    assert rc == 1                      # solve() returned false
    assert n.value > 0 and not phi[:n.value].any()  # and nothing was computed on the host
