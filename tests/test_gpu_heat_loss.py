"""ITCON=2 embedded heat loss on the device against the oracle (bit-exact).

The heat-loss stage kernels (golden/fortran/3dmhdsub.f:1797-1822) are
stepped through the C ABI and every virtual rank's arrays are compared with
the oracle's, ghosts included: hydro with the sink across a z-rank boundary
and partial y tiles, hydro with the TC/QFH ramp (the amplitude follows TIMT
step by step), and MHD.
"""
import sys

import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

HEAT = ["--itcon", "2", "--tp", "1.5", "--xp", "10.2", "--yp", "9.9", "--zp", "27.0",
        "--hh", "6.0"]
CASES = {
    "hydro": ["--npx", "40", "--npy", "30", "--npz", "32", "--npey", "2", "--npez", "2"] + HEAT,
    "ramp": ["--npx", "24", "--npy", "24", "--npz", "24", "--npey", "1", "--npez", "3",
             "--tc", "0.002", "--qfh", "0.001"] + HEAT,
    "mhd": ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2",
            "--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2",
            "--bzp", "10"] + HEAT,
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
def test_heat_loss_matches_oracle(case):
    args = CASES[case]
    say(f"[gpu] ITCON=2 heat loss, {case}: {' '.join(args)}")
    names = ["ru", "rv", "rw", "ro", "tt", "zru", "zrv", "zrw", "zro", "ztt"]
    if "--lmag" in args:
        names += ["bx", "by", "bz", "zbx", "zby", "zbz"]
    run = pkg.Run(args)
    sim = run.make_sim()
    orc = oracle.Sim(args)
    for n in (1, 3):
        sim.step(n)
        orc.step(n)
        compare(sim, orc, names, f"{case} after +{n}")
    if case == "ramp":  # the ramp has passed its midpoint: the loss is near full
        assert orc.scalars()["timt"] > 0.002
    sim.close()


def test_heat_loss_refused_where_not_built():
    args = CASES["hydro"]
    run = pkg.Run(args)
    cfg = run.cfg()
    cfg.variant = 1
    with pytest.raises(pkg.HipError, match="bit-exact variant only"):
        pkg.HipSim(cfg)
