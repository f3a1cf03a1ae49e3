"""Checkpoint and restart on the device (r3d_hip_snapshot_*, --nstart).

* An asynchronous snapshot, with further steps issued while its copy is in
  flight, returns exactly what r3d_hip_download_rank returns at the
  snapshot's point (ghost cells, RK accumulators, scalars).
* Dump (from a snapshot) at step 3, restart with --nstart 3, 4 more steps:
  bit-identical to 7 uninterrupted steps (state and scalars), hydro, MHD,
  and hydro with density diffusion and the ramped ITCON=2 heat loss.
* The HIP path restarted from the reference CPU build's own step-2 dump and
  stepped once reproduces the reference's step-3 dump hashes.
"""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "3dmhd")
MPIRUN = "/opt/conda/bin/mpirun"
X_FIELDS = ["ru", "rv", "rw", "ro", "tt"]
Z_FIELDS = ["zru", "zrv", "zrw", "zro", "ztt"]
HYDRO = ["--npx", "24", "--npy", "24", "--npz", "48", "--npey", "2", "--npez", "2"]
MAG = ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2",
       "--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2", "--bzp", "10",
       "--lrot", "1", "--ry", "0.1", "--ang", "0"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("args", [HYDRO, MAG], ids=["hydro", "mhd"])
def test_snapshot_is_the_download_at_its_point(args):
    say(f"[gpu] snapshot vs download ({'mhd' if '--lmag' in args else 'hydro'})")
    run = pkg.Run(args)
    ref = run.make_sim()
    ref.step(3)
    want = [ref.download_rank(r) for r in range(4)]
    want_sc = ref.scalars()
    ref.close()
    sim = run.make_sim()
    sim.step(3)
    sim.snapshot_begin()
    sim.step(2)  # issued while the snapshot's copy is in flight
    for r in range(4):
        got, sc = sim.snapshot_rank(r)
        for n in want[r]:
            assert same(got[n], want[r][n]), (r, n)
        assert sc == want_sc
    # a second snapshot reuses the buffers
    sim.snapshot_begin()
    got, sc = sim.snapshot_rank(0)
    assert sc["nit"] == 5
    sim.close()


# density diffusion (the dump holds the inverted mirror planes of RO) and the
# TIMT-ramped heat loss (its amplitude must resume from the dumped TIMT)
FEATURES = HYDRO + ["--ngrid", "0", "--id", "1", "--dh", "0.4", "--dv", "0.3", "--itcon", "2",
                    "--tp", "1.5", "--zp", "27", "--hh", "6", "--tc", "0.06", "--qfh", "0.03"]


@pytest.mark.parametrize("args", [HYDRO, MAG, FEATURES], ids=["hydro", "mhd", "features"])
def test_restart_continues_bit_for_bit(args, tmp_path):
    tag = "mhd" if "--lmag" in args else ("features" if "--id" in args else "hydro")
    say(f"[gpu] dump at step 3, restart, 4 more steps ({tag})")
    run = pkg.Run(args)
    names = X_FIELDS + (["bx", "by", "bz"] if "--lmag" in args else [])
    sim = run.make_sim()
    sim.step(3)
    t = run.dump_async(sim, str(tmp_path))
    sim.step(4)  # runs while the dump is written
    t.join()
    assert not t.errors, t.errors
    want = sim.download_global(names)
    want_sc = sim.scalars()
    sim.close()
    assert sorted(os.listdir(tmp_path)) == [f"gstate.step.{r:03d}.0003" for r in range(4)]
    rrun = pkg.Run(args + ["--nstart", "3", "--dumpdir", str(tmp_path)])
    assert rrun.scalars()["nit"] == 3
    rsim = rrun.make_sim()
    rsim.step(4)
    got = rsim.download_global(names)
    got_sc = rsim.scalars()
    rsim.close()
    for n in names:
        assert same(got[n], want[n]), n
    assert got_sc == want_sc


def test_restart_from_the_reference_dump(golden, tmp_path):
    if not (os.path.exists(REF) and os.path.exists(MPIRUN)):
        pytest.skip("reference build (oracle/_ref) or mpirun not available")
    say("[gpu] restart from the reference's step-2 dump (g16 2x2), one step")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([MPIRUN, "-np", "4", REF, "--npx", "16", "--npy", "16", "--npz", "32",
                    "--npey", "2", "--ntotal", "2", "--nstep0", "2"], cwd=tmp_path, env=env,
                   check=True, capture_output=True, timeout=120)
    run = pkg.Run(["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2",
                   "--nstart", "2", "--dumpdir", str(tmp_path)])
    sim = run.make_sim()
    sim.step(1)
    rec = golden["g16_2x2"]["dumps"]["3"]
    bad = []
    for r, want in enumerate(rec):
        got = sim.download_rank(r, X_FIELDS + Z_FIELDS)
        for n in X_FIELDS + Z_FIELDS:
            if hashlib.sha256(got[n].tobytes()).hexdigest()[:20] != want["fields"][n]:
                bad.append((r, n))
    sc = sim.scalars()
    for k in ("dt", "timc", "timt", "umach"):
        if sc[k].hex() != rec[0]["scalars"][k]:
            bad.append((k, sc[k]))
    assert sc["nit"] == 3
    sim.close()
    assert not bad, bad
