"""Restart from gstate dumps (--nstart / --dumpdir, r3d_host_read_gstate).

The reference writes gstate dumps (src/main.cpp:44-93) but cannot restart from
them (src/main.cpp:148-153 stops with "restart (NSTART!=0) is not yet
ported").  These CPU checks pin the host side of the restart on the
reference build's own dumps (oracle/_ref/3dmhd, the geometry of the
reference's golden tests, 16x16x32 on 2x2 ranks):

* the global state and scalars a --nstart 2 run assembles equal, cell for
  cell and bit for bit, every rank's step-2 dump (ghosts included);
* that state, stepped once by the CPU oracle, reproduces the reference's own
  step-3 dump hashes (tests/golden/ref_hashes.json), so nothing a step needs
  is lost in a restart;
* malformed or mismatched dumps are refused with a message.
"""
import hashlib
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import load_package
import gstate
import oracle

pkg = load_package()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "3dmhd")
MPIRUN = "/opt/conda/bin/mpirun"
G16 = ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2"]
PRIM = ["ru", "rv", "rw", "ro", "tt"]


@pytest.fixture(scope="module")
def ref_dumps(tmp_path_factory):
    """The reference build's dumps of steps 1 and 2 (g16, 2x2 ranks)."""
    if not (os.path.exists(REF) and os.path.exists(MPIRUN)):
        pytest.skip("reference build (oracle/_ref) or mpirun not available")
    d = tmp_path_factory.mktemp("refdump")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    # the reference takes npez from the rank count: 4 ranks / npey 2
    args = G16[:6] + ["--npey", "2", "--ntotal", "2", "--nstep0", "1"]
    subprocess.run([MPIRUN, "-np", "4", REF] + args, cwd=d, env=env, check=True,
                   capture_output=True)
    return str(d)


def local_box(run, rank):
    nry, nrz = run.npy // run.npey, run.npz // run.npez
    py, pz = rank % run.npey, rank // run.npey
    return (slice(pz * nrz, pz * nrz + nrz + 2), slice(py * nry, py * nry + nry + 2), slice(None))


def test_restart_state_is_the_reference_dump(ref_dumps):
    run = pkg.Run(G16 + ["--nstart", "2", "--dumpdir", ref_dumps])
    st = run.state()
    sc = run.scalars()
    for r in range(4):
        g = gstate.read(os.path.join(ref_dumps, f"gstate.step.{r:03d}.0002"))
        assert g["header"]["nit"] == 2 and sc["nit"] == 2
        for k in ("dt", "timc", "timt", "timi", "umach"):
            assert sc[k] == g["scalars"][k], (k, sc[k], g["scalars"][k])
        box = local_box(run, r)
        for n in PRIM:
            got, want = st[n][box], g["fields"][n]
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (r, n)
        # the reader returns the same arrays, RK accumulators included
        f, s = run.read_gstate(r, os.path.join(ref_dumps, f"gstate.step.{r:03d}.0002"))
        for n in PRIM + ["zru", "zrv", "zrw", "zro", "ztt"]:
            assert np.array_equal(f[n], g["fields"][n]), (r, n)
        assert s == sc


def test_restarted_state_steps_to_the_reference_step_3(ref_dumps, golden):
    """Restart at step 2 + one oracle step == the reference's step-3 dump."""
    run = pkg.Run(G16 + ["--nstart", "2", "--dumpdir", ref_dumps])
    st, sc = run.state(), run.scalars()
    sim = oracle.Sim(G16)
    for r in range(4):
        box = local_box(run, r)
        for n in PRIM:
            sim.field(r, n)[...] = st[n][box]
    osc = sim.scalars()
    sim.set_scalars(dict(sc, tb=osc["tb"]))
    sim.step(1)
    rec = golden["g16_2x2"]["dumps"]["3"]
    for r, want in enumerate(rec):
        for n in PRIM + ["zru", "zrv", "zrw", "zro", "ztt"]:
            got = oracle_digest(sim.field(r, n))
            assert got == want["fields"][n], (r, n)
        got_sc = sim.scalars()
        for k, hx in want["scalars"].items():
            assert got_sc[k].hex() == hx, (r, k)
        assert int(got_sc["nit"]) == want["nit"]


def oracle_digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:20]


def test_restart_refuses_bad_dumps(ref_dumps, tmp_path):
    src = os.path.join(ref_dumps, "gstate.step.000.0002")
    # a dump of another grid
    with pytest.raises(pkg.HipError, match="nx"):
        pkg.Run(["--npx", "24", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2",
                 "--nstart", "2", "--dumpdir", ref_dumps])
    # missing step
    with pytest.raises(pkg.HipError, match="cannot open"):
        pkg.Run(G16 + ["--nstart", "7", "--dumpdir", ref_dumps])
    # truncated and corrupted copies
    bad = tmp_path / "bad"
    bad.mkdir()
    for r in range(4):
        shutil.copy(os.path.join(ref_dumps, f"gstate.step.{r:03d}.0002"), bad)
    data = open(src, "rb").read()
    open(bad / "gstate.step.000.0002", "wb").write(data[: len(data) // 2])
    with pytest.raises(pkg.HipError, match="truncated"):
        pkg.Run(G16 + ["--nstart", "2", "--dumpdir", str(bad)])
    open(bad / "gstate.step.000.0002", "wb").write(struct.pack("<i", 7) + data[4:])
    with pytest.raises(pkg.HipError, match="not a gstate dump"):
        pkg.Run(G16 + ["--nstart", "2", "--dumpdir", str(bad)])
    # a rank's dump from another step
    open(bad / "gstate.step.000.0002", "wb").write(data)
    shutil.copy(os.path.join(ref_dumps, "gstate.step.001.0001"), bad / "gstate.step.001.0002")
    with pytest.raises(pkg.HipError, match="step 1"):
        pkg.Run(G16 + ["--nstart", "2", "--dumpdir", str(bad)])
