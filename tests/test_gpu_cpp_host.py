"""The C++ host side (cellweave_amd/host, the ScalarField and coupling
interface of crates/engine/src/{pde,coupling}.rs over the C ABI): the
reference's own tests restated in C++ (tests/cpp/test_reference.cpp), run on
the GPU. The binary is built by __graft_entry__.build()."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "test_reference")

pytestmark = pytest.mark.gpu


def test_reference_tests_in_cpp_pass_on_the_gpu():
    assert os.path.exists(BIN), f"{BIN} is missing: run __graft_entry__.build()"
    print("case C++ host side: tests/cpp/build/test_reference", flush=True)
    run = subprocess.run([BIN], capture_output=True, text=True, timeout=600)
    print(run.stdout[-4000:], flush=True)
    assert run.returncode == 0, run.stdout[-4000:] + run.stderr[-2000:]
    assert "FAIL" not in run.stdout
