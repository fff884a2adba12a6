"""The C++ drop-in header (include/rsbd8_hip.hpp) on the GPU: the reference's portable known-answer
tests (multiplatform/main.cpp:10-304) and its documented direction example, compiled against the
header and run as a program (radixsortbidi8_amd/examples/kat_main.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "radixsortbidi8_amd", "lib", "kat_main")


def test_cpp_dropin_known_answers():
    assert os.path.exists(EXE), "build the example with radixsortbidi8_amd/examples/build.sh"
    print("running", EXE, flush=True)
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ok (0 failures)" in p.stdout
