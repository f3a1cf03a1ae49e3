"""The reference-side shim in INTEGRATION.md compiles against the reference's
own headers (include/gpu.hpp, params.hpp, ...) and this build's C ABI headers.
Skipped where the reference sources or MPICH's compiler wrapper are absent."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/include"
MPICXX = "/opt/conda/bin/mpicxx"


def test_shim_compiles_against_the_reference_headers(tmp_path):
    if not (os.path.isdir(REF_INC) and os.path.exists(MPICXX)):
        pytest.skip("reference headers or mpicxx not available")
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```(?:cpp|c\+\+)\n(.*?)```", text, re.S)
    assert len(blocks) == 1
    src = tmp_path / "shim.cpp"
    src.write_text(blocks[0])
    r = subprocess.run([MPICXX, "-std=c++17", "-fsyntax-only", "-I" + REF_INC,
                        "-I" + os.path.join(ROOT, "include"), str(src)],
                       env=dict(os.environ, MPICH_CXX="g++"), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
