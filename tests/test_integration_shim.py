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


BIN = os.path.join(ROOT, "oracle", "_ref", "3dmhd_gpu")
MPIRUN = "/opt/conda/bin/mpirun"


def _driver(tmp_path, backend):
    if not (os.path.exists(BIN) and os.path.exists(MPIRUN)):
        pytest.skip("the linked reference driver is not built (needs the reference sources)")
    return subprocess.run([MPIRUN, "-np", "2", BIN, "--npx", "16", "--npy", "16", "--npz", "32",
                           "--npey", "1", "--ntotal", "1", "--nstep0", "1", "--backend", backend],
                          cwd=tmp_path, capture_output=True, text=True, timeout=120,
                          env=dict(os.environ, OMP_NUM_THREADS="1"))


def test_linked_reference_driver_keeps_its_cpu_backend(tmp_path):
    """The reference driver linked with the shim and the library (oracle/ref.mk)
    still runs its own CPU path unchanged with --backend cpu."""
    r = _driver(tmp_path, "cpu")
    assert r.returncode == 0, r.stderr[-1000:]
    assert os.path.exists(tmp_path / "gstate.step.001.0001")


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present here")
def test_linked_reference_driver_reports_library_errors(tmp_path):
    """Without a GPU, --backend gpu reaches the library, whose error comes back
    as a code; the shim prints it and aborts the MPI job like the reference."""
    r = _driver(tmp_path, "gpu")
    assert r.returncode != 0
    assert "gfx950 backend" in r.stderr and "hipSetDevice" in r.stderr, r.stderr[-1000:]
