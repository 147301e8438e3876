"""The C++ drop-in header (include/cufhedb_amd/cuhedb.hpp) driven like a
cuFHEDB caller (both programs built by `make dropin`):

* tests/cpp/dropin_demo.cpp: CuFFT_Torus<P>::LoadBK /
  GateBootstrappingTLWE2TLWEFFT_st / ifft_st / fft_st and cuHEDB::HomAND /
  HomOR / HomNOT on an explicit Runtime, the coalescing scheduler and the
  reference's key files;
* tests/cpp/dropin_refcalls.cpp: the TPC-H Q1 filter steps written with the
  reference's own signatures (EvalKey argument, stream_id, the global
  cufftlvl1, worker threads with one stream_id each), decrypted results
  against the plaintext predicates."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("prog", ["dropin_demo", "dropin_refcalls"])
def test_cpp_dropin_program(prog):
    exe = os.path.join(ROOT, "_build", prog)
    assert os.path.exists(exe), "build it with `make dropin`"
    print(f"running {prog}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ok" in r.stdout
