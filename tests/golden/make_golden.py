#!/usr/bin/env python3
"""Generate the reference fixtures under tests/golden/ from the reference CPU build.

Sets (python tests/golden/make_golden.py [set ...], default: all):

base -> ref_hashes.json: oracle/_ref/3dmhd (the reference CPU path, built
    from its sources by oracle/ref.mk) under MPICH's mpirun for a set of small
    configurations; for every gstate dump a SHA-256 of each field's raw bytes
    (whole rank array, ghosts included), of each 1-D metric, and the exact
    scalars.
lmag -> ref_lmag.json: what the reference computes with LMAG before its
    FLUXES stops (src/fluxes.cpp:539-541).  "start" records hash the
    gstate.start dumps of oracle/_ref/3dmhd (STATIC with the magnetic layer
    and its density correction, src/static_ic.cpp:94-108, written by
    src/main.cpp:199 before the first step).  "stepdt" records hash what
    STEP's time-step part leaves (oracle/_ref/stepdt, oracle/ref_stepdt.cpp):
    the interior of UU (fast speed), VV (resistive limit), WW1..WW3 and of the
    input fields, with the exact dt and umach (src/step.cpp:121-229).
prod -> ref_prod.json: the production tiling: 504x504x640 hydro on 4x2
    ranks, 3 steps, hashes of the primaries and RK accumulators of every
    rank (no scratch arrays: a fused GPU step does not materialise them).

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


STEPDT = os.path.join(ROOT, "oracle", "_ref", "stepdt")
MHD = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2", "--bzp", "10",
       "--lrot", "1", "--ry", "0.1", "--ang", "0"]

# LMAG start dumps: name -> (args, npey, npez)
LMAG_START = {
    # the rotating-MHD flags at the reference's golden geometry
    "m16_2x4": (["--npx", "16", "--npy", "16", "--npz", "32"] + MHD, 2, 4),
    # the same flags at the rotating-MHD workload size (252x252x512 on 2x4)
    "m252_2x4": (["--npx", "252", "--npy", "252", "--npz", "512"] + MHD, 2, 4),
}

# STEP's limits: name -> (args, npey, npez, flow)
STEPDT_CASES = {
    "sdt_m16_rest": (["--npx", "16", "--npy", "16", "--npz", "32"] + MHD, 2, 4, 0),
    "sdt_m16_flow": (["--npx", "16", "--npy", "16", "--npz", "32"] + MHD, 2, 4, 1),
    # the resistive limit VV = 0.5 dd^2 Rm binds
    "sdt_m16_rm": (["--npx", "16", "--npy", "16", "--npz", "32"] + MHD + ["--rm", "0.01"],
                   2, 4, 1),
    # a strong field: b2 dominates the fast speed
    "sdt_m16_beta": (["--npx", "16", "--npy", "16", "--npz", "32"] + MHD + ["--beta", "0.1"],
                     2, 4, 1),
    # with Re*Pr = 50 the advective limit dd/sqrt(fast speed) binds: dt itself
    # then depends on b2 and on the 2 sqrt(UU (gamma T + b2)) term
    "sdt_m16_adv_rest": (["--npx", "16", "--npy", "16", "--npz", "32"] + MHD +
                         ["--re", "50", "--pr", "1"], 2, 4, 0),
    "sdt_m16_adv_flow": (["--npx", "16", "--npy", "16", "--npz", "32"] + MHD +
                         ["--re", "50", "--pr", "1"], 2, 4, 1),
    "sdt_m16_adv_beta": (["--npx", "16", "--npy", "16", "--npz", "32"] + MHD +
                         ["--re", "50", "--pr", "1", "--beta", "0.1"], 2, 4, 1),
    "sdt_m24_flow_1x2": (["--npx", "24", "--npy", "24", "--npz", "48"] + MHD, 1, 2, 1),
    "sdt_h24_flow": (["--npx", "24", "--npy", "24", "--npz", "48"], 2, 2, 1),
    "sdt_m252_flow": (["--npx", "252", "--npy", "252", "--npz", "512"] + MHD, 2, 4, 1),
}
STEPDT_FIELDS = ["ru", "rv", "rw", "ro", "tt", "bx", "by", "bz", "uu", "vv", "ww1", "ww2", "ww3"]

PROD = {
    # 504x504x640 on 4x2: the library picks 128-plane z chunks, eight 64-wide x
    # tiles (the last 56 wide) and 63 eight-row y tiles, the production tiling
    "p504x640_4x2": (["--npx", "504", "--npy", "504", "--npz", "640"], 4, 2, [3]),
}
PROD_FIELDS = ["ru", "rv", "rw", "ro", "tt", "zru", "zrv", "zrw", "zro", "ztt"]


def run_lmag_start(args, npey, npez):
    """gstate.start of the reference with LMAG; its first step stops in FLUXES."""
    tmp = tempfile.mkdtemp(prefix="r3d_golden_")
    try:
        full = args + ["--npey", str(npey), "--ntotal", "1", "--nstep0", "1"]
        env = dict(os.environ, OMP_NUM_THREADS="1")
        p = subprocess.run([MPIRUN, "-np", str(npey * npez), REF] + full, cwd=tmp, env=env,
                           capture_output=True, text=True)
        if "unsupported feature enabled: LMAG" not in p.stderr:
            raise RuntimeError(f"expected the LMAG stop in FLUXES: rc={p.returncode} "
                               f"{p.stderr[-400:]}")
        return {"args": args, "npey": npey, "npez": npez, "dumps": {"0": [
            hash_dump(os.path.join(tmp, f"gstate.start.{r:03d}.0000"), npey)
            for r in range(npey * npez)]}}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def run_stepdt(args, npey, npez, flow):
    import numpy as np
    import struct
    tmp = tempfile.mkdtemp(prefix="r3d_golden_")
    try:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([MPIRUN, "-np", str(npey * npez), STEPDT, tmp, str(flow)] + args +
                       ["--npey", str(npey)], cwd=tmp, env=env, check=True, capture_output=True)
        ranks = []
        for r in range(npey * npez):
            with open(os.path.join(tmp, f"stepdt.{r}.bin"), "rb") as f:
                nx, ny, nz = struct.unpack("3i", f.read(12))
                dt, umach = struct.unpack("2d", f.read(16))
                rec = {"dt": dt.hex(), "umach": umach.hex(), "interior": {}}
                for name in STEPDT_FIELDS:
                    a = np.fromfile(f, dtype=np.float64, count=nx * ny * nz).reshape(nz, ny, nx)
                    rec["interior"][name] = digest(np.ascontiguousarray(a[1:-1, 1:-1, 1:-1]))
                ranks.append(rec)
        return {"args": args, "npey": npey, "npez": npez, "flow": flow, "ranks": ranks}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def run_prod(args, npey, npez, steps):
    """Like run(), but the dumps are hashed (primaries and RK accumulators)
    and deleted one at a time: the 504x504x640 dumps are 4.5 GB per rank."""
    tmp = tempfile.mkdtemp(prefix="r3d_golden_")
    try:
        nmax = max(steps)
        full = args + ["--npey", str(npey), "--ntotal", str(nmax), "--nstep0", str(nstep0(steps))]
        env = dict(os.environ, OMP_NUM_THREADS=str(max(1, 8 // (npey * npez))))
        subprocess.run([MPIRUN, "-np", str(npey * npez), REF] + full, cwd=tmp, env=env,
                       check=True, capture_output=True)
        rec = {"args": args, "npey": npey, "npez": npez, "dumps": {}}
        for st in steps:
            out = []
            for r in range(npey * npez):
                path = os.path.join(tmp, f"gstate.step.{r:03d}.{st:04d}")
                g = gstate.read(path)
                out.append({"fields": {k: digest(g["fields"][k]) for k in PROD_FIELDS},
                            "scalars": {k: g["scalars"][k].hex()
                                        for k in ("dt", "timc", "timt", "timi", "umach")},
                            "nit": g["header"]["nit"]})
                del g
                os.unlink(path)
            rec["dumps"][str(st)] = out
        return rec
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def write(name, out):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main():
    sets = sys.argv[1:] or ["base", "lmag", "prod"]
    if not os.path.exists(REF):
        sys.exit(f"reference build missing: {REF} (make -f oracle/ref.mk)")
    if "base" in sets:
        out = {}
        for name, (args, npey, npez, steps) in CONFIGS.items():
            print(name, flush=True)
            out[name] = run(name, args, npey, npez, steps)
        write("ref_hashes.json", out)
    if "lmag" in sets:
        out = {"start": {}, "stepdt": {}}
        for name, (args, npey, npez) in LMAG_START.items():
            print(name, flush=True)
            out["start"][name] = run_lmag_start(args, npey, npez)
        for name, (args, npey, npez, flow) in STEPDT_CASES.items():
            print(name, flush=True)
            out["stepdt"][name] = run_stepdt(args, npey, npez, flow)
        write("ref_lmag.json", out)
    if "prod" in sets:
        out = {}
        for name, (args, npey, npez, steps) in PROD.items():
            print(name, flush=True)
            out[name] = run_prod(args, npey, npez, steps)
        write("ref_prod.json", out)


if __name__ == "__main__":
    main()
