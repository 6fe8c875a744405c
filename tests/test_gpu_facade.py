"""The C++ PICCycle facade (scdat_amd/facade/PICCycle.hpp) running the
reference's PICCycle boundary tests (PICCycleBoundary_test.cpp:121-221) on the
device; the program is built by __graft_entry__.build() / tests/cpp/Makefile."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "picycle_facade_test")


@pytest.mark.gpu
def test_cpp_facade_reference_boundary_tests():
    assert os.path.exists(EXE), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "all facade checks passed" in r.stdout
