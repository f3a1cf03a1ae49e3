"""The drop-in, end to end: the reference's own driver runs this library through
its `--backend gpu` loop.

oracle/_ref/3dmhd_gpu is the reference's src/main.cpp (its GPU time loop,
src/main.cpp:234-275, compiled with R3D_HAVE_GPU), linked with the
INTEGRATION.md shim (the include/gpu.hpp surface over include/r3d_hip.h) and
librast3dmhd_hip.so; oracle/ref.mk builds it where the reference sources are.
Launched like the reference (mpirun, one rank per subdomain, `--backend gpu`),
it must write the gstate dumps of the reference CPU build
(tests/golden/ref_hashes.json) bit for bit: every rank's primaries with their
ghosts, the RK accumulators, the metrics and the scalars.  The one exception
is umach: the reference's GPU loop does not copy it into the dumped state
(src/main.cpp:259-263), so its dumps keep the value of the initial setup.
"""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import gstate
from conftest import ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "oracle", "_ref", "3dmhd_gpu")
MPIRUN = "/opt/conda/bin/mpirun"
FIELDS = ["ru", "rv", "rw", "ro", "tt", "zru", "zrv", "zrw", "zro", "ztt"]
Y_METRICS = ("wyy", "dyydy", "d2yydy2", "ddy")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:20]


def run_driver(tmp, args, npey, npez, ntotal, nstep0):
    if not os.path.exists(BIN):
        pytest.skip(f"{BIN} not built (needs the reference sources: make -f oracle/ref.mk)")
    cmd = [MPIRUN, "-np", str(npey * npez), BIN] + args + [
        "--npey", str(npey), "--ntotal", str(ntotal), "--nstep0", str(nstep0),
        "--backend", "gpu", "--gpu", "0"]
    say(f"[gpu] reference driver: {' '.join(cmd[3:])}")
    p = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True, timeout=240,
                       env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert p.returncode == 0, p.stderr[-2000:]
    return p.stderr


def compare(tmp, want, st, npey):
    tag = "start" if st == 0 else "step"
    bad = []
    for r, rec in enumerate(want):
        g = gstate.read(os.path.join(tmp, f"gstate.{tag}.{r:03d}.{st:04d}"))
        for n in FIELDS:
            if digest(g["fields"][n]) != rec["fields"][n]:
                bad.append((r, n))
        for m, h in rec["metrics"].items():
            v = g["metrics"][m]
            if npey == 1 and m in Y_METRICS:
                v = v[: g["header"]["ny"] - 1]
            if digest(v) != h:
                bad.append((r, "metric:" + m))
        for k in ("dt", "timc", "timt", "timi"):
            if g["scalars"][k].hex() != rec["scalars"][k]:
                bad.append((r, k))
        if g["header"]["nit"] != rec["nit"]:
            bad.append((r, "nit"))
    return bad


@pytest.mark.parametrize("name,ntotal,nstep0", [("h64_1x2", 2, 1), ("h64_1x2", 100, 100),
                                                ("g16_2x2", 3, 1), ("b_rot", 4, 4)])
def test_reference_driver_with_the_library_matches_reference_dumps(golden, tmp_path, name,
                                                                    ntotal, nstep0):
    cfg = golden[name]
    run_driver(str(tmp_path), cfg["args"], cfg["npey"], cfg["npez"], ntotal, nstep0)
    checked = 0
    for st in sorted(int(s) for s in cfg["dumps"]):
        if st > ntotal or (st != 0 and st % nstep0 and st != ntotal):
            continue
        bad = compare(str(tmp_path), cfg["dumps"][str(st)], st, cfg["npey"])
        assert not bad, f"{name} step {st}: {bad[:8]}"
        checked += 1
    assert checked >= 2
