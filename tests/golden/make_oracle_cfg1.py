#!/usr/bin/env python3
"""Generate tests/golden/oracle_cfg1.json: the rotating-MHD workload from the CPU oracle.

The reference cannot step with LMAG (its FLUXES stops, src/fluxes.cpp:539-541),
so the full-size rotating-MHD run (252x252x512 on 2x4 virtual ranks,
--lmag 1 --rm 50 --beta 10 --ampb 1 --bfh 2 --bzp 10 --lrot 1 --ry 0.1 --ang 0,
the reference's ran2 seed) is taken from the oracle (oracle/r3d_oracle.c,
whose LMAG set-up and time-step limits are pinned to the reference build and
whose LMAG fluxes are pinned by closed forms).  After steps 1, 20 and 200 it
records, per virtual rank, a SHA-256 of the primaries (whole rank array,
ghosts included) and of the RK accumulators' interior, with the exact
scalars.  The GPU test tests/test_gpu_production.py compares the HIP path
at the same size against them.

Runs in the development container (about 40 minutes on 8 threads).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

MHD = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2", "--bzp", "10",
       "--lrot", "1", "--ry", "0.1", "--ang", "0"]
ARGS = ["--npx", "252", "--npy", "252", "--npz", "512"] + MHD
NPEY, NPEZ = 2, 4
STEPS = [1, 20, 200]
X = ["ru", "rv", "rw", "ro", "tt", "bx", "by", "bz"]
Z = ["zru", "zrv", "zrw", "zro", "ztt", "zbx", "zby", "zbz"]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:20]


def main():
    sim = oracle.Sim(ARGS + ["--npey", str(NPEY), "--npez", str(NPEZ)])
    out = {"args": ARGS, "npey": NPEY, "npez": NPEZ, "dumps": {}}
    done = 0
    t0 = time.time()
    for st in STEPS:
        for _ in range(st - done):
            sim.step(1)
            done += 1
            if done % 10 == 0:
                print(f"step {done} ({time.time() - t0:.0f} s)", flush=True)
        ranks = []
        for r in range(sim.nranks):
            rec = {n: digest(sim.field(r, n)) for n in X}
            rec.update({n: digest(sim.field(r, n)[1:-1, 1:-1, 1:-1]) for n in Z})
            ranks.append(rec)
        sc = sim.scalars()
        out["dumps"][str(st)] = {"ranks": ranks,
                                 "scalars": {k: sc[k].hex() for k in ("dt", "timc", "timt",
                                                                      "umach")},
                                 "nit": int(sc["nit"])}
        with open(os.path.join(HERE, "oracle_cfg1.json"), "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
        print(f"step {st} recorded", flush=True)


if __name__ == "__main__":
    main()
