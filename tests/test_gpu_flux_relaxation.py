"""Flux relaxation (RLAX, ITC=0 / IZC=1) on the device against the oracle.

Per stage the device forms TTM and FCONM with the reference's summation
order, updates TFAC from the bottom-wall mean and builds RLAX*TFAC*CORRECT
per plane in one thread (golden/fortran/3dmhdsub.f:1534-1547, 1625-1771),
and the DEN stage kernels add it to FT.  Every virtual rank's arrays are
compared with the oracle's, ghosts included, bit for bit.
"""
import sys

import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

RLX = ["--ngrid", "0", "--itcon", "0", "--izcon", "1", "--rlax", "0.5", "--toff", "0.5"]
CASES = {
    "hydro": ["--npx", "40", "--npy", "30", "--npz", "32", "--npey", "2", "--npez", "2",
              "--tstart", "-1"] + RLX,
    "late_start": ["--npx", "24", "--npy", "24", "--npz", "24", "--npey", "1", "--npez", "3",
                   "--tstart", "0.03"] + RLX,
    "ragged": ["--npx", "70", "--npy", "18", "--npz", "30", "--npey", "3", "--npez", "3",
               "--tstart", "-1"] + RLX,
    "with_density": ["--npx", "40", "--npy", "30", "--npz", "32", "--npey", "2", "--npez", "2",
                     "--tstart", "-1", "--id", "1", "--dh", "0.4", "--dv", "0.3"] + RLX,
    "mhd": ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2",
            "--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2",
            "--bzp", "10", "--tstart", "-1"] + RLX,
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare(sim, orc, names, tag):
    bad = []
    for r in range(orc.nranks):
        got = sim.download_rank(r, names)
        for n in names:
            g, w = got[n], orc.field(r, n)
            if n.startswith("z"):
                g, w = g[1:-1, 1:-1, 1:-1], w[1:-1, 1:-1, 1:-1]
            if not np.array_equal(g.view(np.uint64), w.view(np.uint64)):
                bad.append((r, n, int((g != w).sum()), float(np.abs(g - w).max())))
    so, sg = orc.scalars(), sim.scalars()
    for k in ("dt", "timc", "timt", "umach", "nit"):
        if so[k] != sg[k]:
            bad.append((k, sg[k], so[k]))
    assert not bad, f"{tag}: {bad[:6]}"


@pytest.mark.parametrize("case", list(CASES))
def test_flux_relaxation_matches_oracle(case):
    args = CASES[case]
    say(f"[gpu] flux relaxation, {case}: {' '.join(args)}")
    names = ["ru", "rv", "rw", "ro", "tt", "zru", "zrv", "zrw", "zro", "ztt"]
    if "--lmag" in args:
        names += ["bx", "by", "bz", "zbx", "zby", "zbz"]
    sim = pkg.Run(args).make_sim()
    orc = oracle.Sim(args)
    for n in (1, 3):
        sim.step(n)
        orc.step(n)
        compare(sim, orc, names, f"{case} after +{n}")
    if case == "hydro":
        sim.step(1, 2)
        orc.step(1, 2)
        compare(sim, orc, names, f"{case} after a 2-stage step")
    st = orc.relax_state()
    assert st["tfac"] != 1.0  # TFAC moved
    sim.close()


def test_flux_relaxation_changes_the_step():
    grid = ["--npx", "40", "--npy", "30", "--npz", "32", "--npey", "2", "--npez", "2",
            "--ngrid", "0", "--itcon", "0", "--izcon", "1"]
    say("[gpu] flux relaxation on vs off")
    a = pkg.Run(grid + ["--rlax", "0.5", "--toff", "0.5", "--tstart", "-1"]).make_sim()
    b = pkg.Run(grid).make_sim()
    a.step(2)
    b.step(2)
    assert not np.array_equal(a.download_global(["tt"])["tt"], b.download_global(["tt"])["tt"])
    a.close()
    b.close()
