"""The C++ host side and its tests compile (syntax and types only, no GPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
@pytest.mark.parametrize("src", ["cellweave_amd/host/cellweave.cpp", "tests/cpp/test_reference.cpp"])
def test_host_sources_compile(src):
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                    "-I", os.path.join(ROOT, "cellweave_amd", "host"), os.path.join(ROOT, src)],
                   check=True, timeout=120)
