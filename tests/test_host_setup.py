"""The product's host setup (librast3dmhd_host) against the oracle.

The global initial state and metrics built by the product must equal, bit for
bit, what the oracle (pinned to the reference dumps) holds in every virtual
rank, ghost cells included.
"""
import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()

CASES = [
    ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2"],
    ["--npx", "24", "--npy", "24", "--npz", "48", "--npey", "1", "--npez", "2", "--lrot", "1",
     "--ry", "0.1"],
    ["--npx", "40", "--npy", "24", "--npz", "48", "--npey", "4", "--npez", "3", "--ngrid", "2"],
    ["--npx", "24", "--npy", "24", "--npz", "48", "--npey", "2", "--npez", "2", "--pzp", "20",
     "--sigma", "3", "--polys", "2", "--seed", "-1234"],
    ["--npx", "16", "--npy", "32", "--npz", "64", "--npey", "2", "--npez", "4", "--lmag", "1",
     "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2", "--bzp", "10"],
    # LREM stratification (its conductivity replaces the PZP profile)
    ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2", "--lrem", "1",
     "--ngrid", "0", "--zmax", "4", "--pzp", "2"],
    ["--npx", "16", "--npy", "24", "--npz", "48", "--npey", "3", "--npez", "3", "--lrem", "1",
     "--ngrid", "0", "--zmax", "3", "--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1",
     "--bfh", "0.5", "--bzp", "1.5", "--theta", "0.3"],
]


def rank_view(glob, npey, npez, rank):
    py, pz = rank % npey, rank // npey
    Z, Y, X = glob.shape
    nry, nrz = (Y - 2) // npey, (Z - 2) // npez
    return glob[pz * nrz:pz * nrz + nrz + 2, py * nry:py * nry + nry + 2, :]


@pytest.mark.parametrize("args", CASES, ids=[str(i) for i in range(len(CASES))])
def test_initial_state_matches_oracle(args):
    names = ("ru", "rv", "rw", "ro", "tt") + (("bx", "by", "bz") if "--lmag" in args else ())
    run = pkg.Run(args)
    sim = oracle.Sim(args)
    st = run.state()
    for r in range(sim.nranks):
        for name in names:
            want = sim.field(r, name)
            got = rank_view(st[name], sim.npey, sim.npez, r)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (r, name)
        rm = run.rank_metrics(r)
        for name in oracle.METRICS:
            want = sim.metric(r, name)
            got = rm[name]
            if sim.npey == 1 and name in ("wyy", "dyydy", "d2yydy2", "ddy"):
                want, got = want[:-1], got[:-1]
            assert np.array_equal(got, want), (r, name)
    assert run.tb() == sim.scalars()["tb"]
    if "--lrem" in args:
        pr = sim.lrem_profile()
        cfg = run.cfg()
        assert cfg.lrem == 1
        assert (cfg.tu, cfg.dztu, cfg.dztb) == (pr["TU"], pr["DZTU"], pr["DZTB"])


def test_cfg_constants_match_oracle():
    args = CASES[1]
    run = pkg.Run(args)
    cfg = run.cfg()
    d = oracle.Sim(args, init=False).derived()
    for k in ("c13", "c23", "c43", "gam1", "gam2", "gam3", "zeta1", "zeta2", "hx", "h2x",
              "hy", "h2y", "hz", "h2z", "repr", "ore", "cv", "ocv", "omx", "omz"):
        assert getattr(cfg, k) == d[k], k
    assert cfg.sff == d["sf"] and cfg.gamma == d["gamma"]


def test_bad_flags_rejected():
    with pytest.raises(pkg.HipError, match="unknown option"):
        pkg.Run(["--nope", "3"])
    with pytest.raises(pkg.HipError, match="not divisible"):
        pkg.Run(["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "3"])
    with pytest.raises(pkg.HipError, match="LREM"):
        pkg.Run(["--npx", "16", "--npy", "16", "--npz", "32", "--lrem", "1"])
