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
Skipped where the reference sources are absent."""
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
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref"), "dropin"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return {k: os.path.join(DROP, k + ".o") for k in ("PICCycle_device", "main_PICcore", "PicMccSurfaceCurrentSampler")}


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
