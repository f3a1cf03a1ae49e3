"""Density diffusion (ID=1/2) on the device against the oracle (bit-exact).

Per stage the device mirrors the outer RO planes, forms the horizontal mean
with the reference's summation order (per y rank, then across y ranks) and
adds the diffusion terms to FR in the DEN stage kernels
(golden/fortran/3dmhdsub.f:1306-1380).  Every virtual rank's arrays are
compared with the oracle's, ghosts included (the outer RO planes hold the
inverted mirror after a stage, as the reference's FLUXES leaves them).
"""
import sys

import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

DEN = ["--ngrid", "0", "--id", "1", "--dh", "0.4", "--dv", "0.3"]
CASES = {
    "id1": ["--npx", "40", "--npy", "30", "--npz", "32", "--npey", "2", "--npez", "2"] + DEN,
    "id2": ["--npx", "24", "--npy", "24", "--npz", "24", "--npey", "1", "--npez", "3",
            "--ngrid", "0", "--id", "2", "--dh", "0.4", "--dv", "0.3", "--zmax", "2.0"],
    "ragged": ["--npx", "70", "--npy", "18", "--npz", "30", "--npey", "3", "--npez", "3"] + DEN,
    "dh_only": ["--npx", "24", "--npy", "16", "--npz", "24", "--npey", "4", "--npez", "2",
                "--ngrid", "0", "--id", "1", "--dh", "0.4"],
    "with_heat": ["--npx", "40", "--npy", "30", "--npz", "32", "--npey", "2", "--npez", "2",
                  "--itcon", "2", "--tp", "1.5", "--zp", "27.0", "--hh", "6.0"] + DEN,
    "mhd": ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2",
            "--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2",
            "--bzp", "10"] + DEN,
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
def test_density_diffusion_matches_oracle(case):
    args = CASES[case]
    say(f"[gpu] density diffusion, {case}: {' '.join(args)}")
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
    if case == "id1":  # a truncated step leaves the mirror of its last stage
        sim.step(1, 2)
        orc.step(1, 2)
        compare(sim, orc, names, f"{case} after a 2-stage step")
    sim.close()


def test_density_diffusion_changes_the_step():
    grid = ["--npx", "40", "--npy", "30", "--npz", "32", "--npey", "2", "--npez", "2"]
    say("[gpu] density diffusion on vs off")
    a = pkg.Run(grid + DEN).make_sim()
    b = pkg.Run(grid + ["--ngrid", "0"]).make_sim()
    a.step(2)
    b.step(2)
    ra, rb = a.download_global(["ro"])["ro"], b.download_global(["ro"])["ro"]
    assert not np.array_equal(ra, rb)
    a.close()
    b.close()


def test_density_diffusion_refused_where_not_built():
    cfg = pkg.Run(CASES["id1"]).cfg()
    cfg.variant = 1
    with pytest.raises(pkg.HipError, match="bit-exact variant only"):
        pkg.HipSim(cfg)


@pytest.mark.parametrize("extra", [DEN, ["--itcon", "2", "--tp", "1.5", "--zp", "27", "--hh", "6"],
                                   ["--ngrid", "0", "--itcon", "0", "--izcon", "1", "--rlax", "0.5",
                                    "--toff", "0.5"],
                                   ["--ngrid", "0", "--lrem", "1", "--zmax", "4"]],
                         ids=["density", "heat", "relax", "lrem"])
def test_extra_physics_refused_by_ensemble_and_adjoint(extra):
    """The extra physics branches run in the single-run bit-exact path only."""
    args = ["--npx", "24", "--npy", "16", "--npz", "24", "--npey", "1", "--npez", "2"] + extra
    run = pkg.Run(args)
    with pytest.raises(pkg.HipError, match="not built in an ensemble"):
        pkg.Ensemble([run, pkg.Run(args)])
    with pytest.raises(pkg.HipError, match="adjoint"):
        pkg.Adjoint(run.cfg(), run.metrics())
