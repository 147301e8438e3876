"""The C++ drop-in header (include/cufhedb_amd/cuhedb.hpp) driven like a
cuFHEDB caller: CuFFT_Torus<P>::LoadBK / GateBootstrappingTLWE2TLWEFFT_st /
ifft_st / fft_st and cuHEDB::HomAND / HomOR / HomNOT (tests/cpp/dropin_demo.cpp,
built by `make dropin`)."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_dropin_program():
    exe = os.path.join(ROOT, "_build", "dropin_demo")
    assert os.path.exists(exe), "build it with `make dropin`"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ok" in r.stdout
