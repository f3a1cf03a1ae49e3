"""LREM on the device against the oracle (bit-exact).

The HIP path's LREM stratification (M. Rempel; golden/fortran/3dmhdsub.f:
437-490, 817-851, 1316, 1534-1596, 2852-2952) must equal the oracle's
restatement (pinned by the closed forms of tests/test_lrem.py) bit for bit:
every virtual rank's primaries with their ghost cells, the RK accumulators'
interior and the scalars, after each checked step.  Per stage the device
mirrors RO's outer planes, forms TTM with the reference's horizontal-mean
order and HRAD per plane, and the DEN stage kernels diffuse TT - TTM with
unit conductivity; the time-step limit takes 1 + RKAPA and the walls TU,
DZTU, DZTB.
"""
import sys

import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

X5 = ["ru", "rv", "rw", "ro", "tt"]
MAG = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "0.5", "--bzp", "1.5"]
LREM = ["--lrem", "1", "--ngrid", "0", "--zmax", "4"]
CASES = {
    "plume": ["--npx", "32", "--npy", "24", "--npz", "48", "--npey", "2", "--npez", "2"] + LREM,
    "const_t_flux_bottom": ["--npx", "32", "--npy", "24", "--npz", "48", "--npey", "2", "--npez",
                            "2", "--tp", "0", "--itcon", "0", "--izcon", "1"] + LREM,
    "flux_top": ["--npx", "40", "--npy", "18", "--npz", "30", "--npey", "3", "--npez", "3", "--tp",
                 "0", "--itcon", "1"] + LREM,
    "mhd": ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2"] + LREM + MAG,
    "with_density_diffusion": ["--npx", "32", "--npy", "24", "--npz", "48", "--npey", "2",
                               "--npez", "2", "--id", "1", "--dh", "0.4", "--dv", "0.3"] + LREM,
    "one_z_rank": ["--npx", "24", "--npy", "24", "--npz", "40", "--npey", "2", "--npez", "1",
                   "--tp", "0", "--itcon", "1", "--izcon", "1"] + LREM,
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
def test_lrem_matches_oracle(case):
    args = CASES[case]
    say(f"[gpu] LREM {case}: {' '.join(args)}")
    run = pkg.Run(args)
    sim = run.make_sim()
    orc = oracle.Sim(args)
    names = X5 + (["bx", "by", "bz"] if "--lmag" in args else [])
    names = names + ["z" + n for n in names]
    for n in (1, 2, 2):
        sim.step(n)
        orc.step(n)
        compare(sim, orc, names, f"{case} +{n}")
    sim.close()


@pytest.mark.parametrize("nstages", [1, 2])
def test_lrem_truncated_steps_match_oracle(nstages):
    args = CASES["const_t_flux_bottom"]
    run = pkg.Run(args)
    sim = run.make_sim()
    orc = oracle.Sim(args)
    sim.step(1)
    orc.step(1)
    sim.step(1, nstages)
    orc.step(1, nstages)
    compare(sim, orc, X5 + ["z" + n for n in X5], f"truncated {nstages}")
    sim.close()


def test_lrem_full_depth_chunks_match_oracle():
    args = ["--npx", "40", "--npy", "16", "--npz", "140", "--npey", "2", "--npez", "2"] + LREM
    run = pkg.Run(args)
    sim = run.make_sim(zchunk=128)
    orc = oracle.Sim(args)
    sim.step(2)
    orc.step(2)
    compare(sim, orc, X5 + ["z" + n for n in X5], "kc=128")
    sim.close()


def test_lrem_refused_by_the_fma_variant():
    run = pkg.Run(CASES["plume"])
    cfg = run.cfg()
    cfg.variant = 1
    with pytest.raises(pkg.HipError, match="bit-exact variant only"):
        pkg.HipSim(cfg)
