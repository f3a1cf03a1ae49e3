#!/usr/bin/env python3
"""Generate tests/golden/ref_hashes.json from the reference CPU build.

Runs oracle/_ref/3dmhd (the reference CPU path, built from its sources by
oracle/ref.mk) under MPICH's mpirun for a set of small configurations and
records, for every gstate dump, a SHA-256 of each field's raw bytes (whole
rank array, ghosts included), of each 1-D metric, and the exact scalars.

The one exception is the last y ghost of the y metrics when npey == 1: the
reference reads one element past its coordinate vector there
(src/grid.cpp:190-192), so that value is unspecified; it is left out of the
metric hash (the time step never reads it).

This script runs only where the reference build exists (the development
container); the JSON it writes is the committed fixture.
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import gstate  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "3dmhd")
MPIRUN = "/opt/conda/bin/mpirun"

# name -> (extra args, npey, npez, steps at which dumps are hashed)
CONFIGS = {
    # the CPU-sized headline case (hydro, 64x64x128, 1x2), 100 steps
    "h64_1x2": (["--npx", "64", "--npy", "64", "--npz", "128"], 1, 2, [1, 2, 100]),
    # the reference's own golden geometry (16x16x32 on 2x2)
    "g16_2x2": (["--npx", "16", "--npy", "16", "--npz", "32"], 2, 2, [1, 2, 3]),
    # decomposition family on one geometry
    "d32_1x2": (["--npx", "32", "--npy", "32", "--npz", "64"], 1, 2, [5]),
    "d32_2x2": (["--npx", "32", "--npy", "32", "--npz", "64"], 2, 2, [5]),
    "d32_4x2": (["--npx", "32", "--npy", "32", "--npz", "64"], 4, 2, [5]),
    "d32_2x4": (["--npx", "32", "--npy", "32", "--npz", "64"], 2, 4, [5]),
    "d32_1x4": (["--npx", "32", "--npy", "32", "--npz", "64"], 1, 4, [5]),
    # boundary / physics branches (24x24x48 on 2x2)
    "b_itcon0": (["--npx", "24", "--npy", "24", "--npz", "48", "--itcon", "0"], 2, 2, [4]),
    "b_tp0": (["--npx", "24", "--npy", "24", "--npz", "48", "--tp", "0"], 2, 2, [4]),
    "b_itcon0_tp0": (["--npx", "24", "--npy", "24", "--npz", "48", "--itcon", "0", "--tp", "0"], 2, 2, [4]),
    "b_itcon4": (["--npx", "24", "--npy", "24", "--npz", "48", "--itcon", "4"], 2, 2, [4]),
    "b_izcon1": (["--npx", "24", "--npy", "24", "--npz", "48", "--izcon", "1"], 2, 2, [4]),
    "b_rot": (["--npx", "24", "--npy", "24", "--npz", "48", "--lrot", "1", "--ry", "0.1",
               "--ang", "0.3"], 2, 2, [4]),
    "b_tc": (["--npx", "24", "--npy", "24", "--npz", "48", "--tc", "0.0005"], 2, 2, [4]),
    "b_ngrid0": (["--npx", "24", "--npy", "24", "--npz", "48", "--ngrid", "0"], 2, 2, [4]),
    "b_ngrid2": (["--npx", "24", "--npy", "24", "--npz", "48", "--ngrid", "2"], 2, 2, [4]),
    "b_pzp": (["--npx", "24", "--npy", "24", "--npz", "48", "--pzp", "20", "--sigma", "3",
               "--polys", "2"], 2, 2, [4]),
    "b_re_theta": (["--npx", "24", "--npy", "24", "--npz", "48", "--re", "3.3", "--theta",
                    "0.15"], 2, 2, [4]),
    # uneven x/y extents and a longer run with rotation
    "u40x24_2x2": (["--npx", "40", "--npy", "24", "--npz", "48", "--lrot", "1", "--ry", "0.5"],
                   2, 2, [20]),
}


def digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()[:20]


def hash_dump(path, npey):
    g = gstate.read(path)
    ny = g["header"]["ny"]
    out = {"fields": {k: digest(v) for k, v in g["fields"].items()}, "metrics": {}}
    for k, v in g["metrics"].items():
        if npey == 1 and k in ("wyy", "dyydy", "d2yydy2", "ddy"):
            v = v[: ny - 1]
        out["metrics"][k] = digest(v)
    out["scalars"] = {k: g["scalars"][k].hex() for k in ("dt", "timc", "timt", "timi", "umach")}
    out["nit"] = g["header"]["nit"]
    return out


def nstep0(steps):
    """Largest dump cadence that still hits every requested step."""
    from math import gcd
    g = 0
    for st in steps:
        g = gcd(g, st)
    return g


def run(name, args, npey, npez, steps):
    tmp = tempfile.mkdtemp(prefix="r3d_golden_")
    try:
        nmax = max(steps)
        full = args + ["--npey", str(npey), "--ntotal", str(nmax), "--nstep0", str(nstep0(steps))]
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([MPIRUN, "-np", str(npey * npez), REF] + full, cwd=tmp, env=env,
                       check=True, capture_output=True)
        rec = {"args": args, "npey": npey, "npez": npez, "dumps": {}}
        for st in [0] + steps:
            tag = "start" if st == 0 else "step"
            rec["dumps"][str(st)] = [
                hash_dump(os.path.join(tmp, f"gstate.{tag}.{r:03d}.{st:04d}"), npey)
                for r in range(npey * npez)]
        return rec
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    if not os.path.exists(REF):
        sys.exit(f"reference build missing: {REF} (make -f oracle/ref.mk)")
    out = {}
    for name, (args, npey, npez, steps) in CONFIGS.items():
        print(name, flush=True)
        out[name] = run(name, args, npey, npez, steps)
    with open(os.path.join(HERE, "ref_hashes.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
