"""Batched ensemble (include/r3d_ens.h): every member bit-identical to the same
run stepped alone through the plain path, whatever the tile shape, with or
without graph replay, and after the autotuner's trial runs (which must leave
the members' state untouched).
"""
import sys

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

X_FIELDS = ["ru", "rv", "rw", "ro", "tt"]
XZ_FIELDS = X_FIELDS + ["zru", "zrv", "zrw", "zro", "ztt"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def single_run(args, nsteps):
    sim = pkg.Run(args).make_sim()
    sim.step(nsteps)
    out = sim.download_global(XZ_FIELDS)
    sc = sim.scalars()
    sim.close()
    return out, sc


def compare_member(ens, m, want, want_sc, tag):
    got = ens.download_global(m, XZ_FIELDS)
    bad = []
    for n in XZ_FIELDS:
        g, w = got[n], want[n]
        if n.startswith("z"):
            g, w = g[1:-1, 1:-1, 1:-1], w[1:-1, 1:-1, 1:-1]
        if not np.array_equal(g.view(np.uint64), w.view(np.uint64)):
            bad.append((n, int((g != w).sum())))
    sc = ens.scalars(m)
    for k in ("dt", "timc", "timt", "umach", "nit"):
        if sc[k] != want_sc[k]:
            bad.append((k, sc[k], want_sc[k]))
    assert not bad, f"{tag} member {m}: {bad[:6]}"


SMALL = [
    ["--npx", "48", "--npy", "48", "--npz", "96", "--npey", "1", "--npez", "2", "--re", "1.5",
     "--theta", "0.12", "--seed", "-5"],
    ["--npx", "80", "--npy", "80", "--npz", "64", "--npey", "2", "--npez", "2", "--re", "3.25",
     "--theta", "0.27", "--seed", "-6"],
    ["--npx", "20", "--npy", "12", "--npz", "30", "--npey", "1", "--npez", "3", "--itcon", "0",
     "--izcon", "1", "--seed", "-7"],
    ["--npx", "64", "--npy", "40", "--npz", "48", "--npey", "2", "--npez", "2", "--lrot", "1",
     "--ry", "0.2", "--ang", "0.3", "--seed", "-8"],
]


@pytest.mark.parametrize("tiles", [(9, 0), (4, 16), (8, 64), (4, 8), (9, 32), (16, 32, 32),
                                   (8, 16, 32), (16, 0, 32), (28, 32, 16), (16, 16, 16)],
                         ids=["ty9_default", "ty4_kc16", "ty8_kc64", "ty4_kc8", "ty9_kc32",
                              "tw32_ty16_kc32", "tw32_ty8_kc16", "tw32_ty16_default",
                              "tw16_ty28_kc32", "tw16_ty16_kc16"])
def test_ensemble_members_match_single_runs(tiles):
    say(f"[gpu] ensemble of {len(SMALL)} uneven members, tiles {tiles}")
    want = [single_run(a, 6) for a in SMALL]
    ens = pkg.Ensemble([pkg.Run(a) for a in SMALL])
    ens.set_tiles(*tiles)
    ens.step(2)
    ens.set_graph(False)  # plain launches in the middle
    ens.step(1)
    ens.set_graph(True)
    ens.step(3)
    for m in range(len(SMALL)):
        compare_member(ens, m, *want[m], f"tiles {tiles}")
    ens.close()


def test_autotune_restores_state_and_keeps_results():
    say("[gpu] ensemble autotune")
    want = [single_run(a, 4) for a in SMALL]
    ens = pkg.Ensemble([pkg.Run(a) for a in SMALL])
    ens.step(1)
    ms = ens.autotune(2)
    ty, kc, tw, nt = ens.tile_shape()
    say(f"[gpu]   chose ty={ty} kc={kc} tw={tw} ({nt} tiles), {ms:.3f} ms/step")
    assert (tw, ty) in ((64, 4), (64, 8), (64, 9), (32, 8), (32, 16), (16, 16), (16, 28))
    assert kc in (8, 16, 32, 64) and ms > 0
    ens.step(3)
    for m in range(len(SMALL)):
        compare_member(ens, m, *want[m], "after autotune")
    ens.close()


def test_ensemble_mhd_members_match_single_runs():
    mag = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2",
           "--bzp", "10", "--lrot", "1", "--ry", "0.1", "--ang", "0"]
    members = [
        ["--npx", "32", "--npy", "32", "--npz", "48", "--npey", "2", "--npez", "2", "--seed",
         "-11"] + mag,
        ["--npx", "70", "--npy", "24", "--npz", "40", "--npey", "1", "--npez", "2", "--seed",
         "-12", "--re", "3.5"] + mag,
    ]
    say("[gpu] MHD ensemble")
    want = [single_run(a, 3) for a in members]
    xs = ["bx", "by", "bz"]
    want_b = []
    for a in members:
        sim = pkg.Run(a).make_sim()
        sim.step(3)
        want_b.append(sim.download_global(xs))
        sim.close()
    ens = pkg.Ensemble([pkg.Run(a) for a in members])
    ens.step(3)
    for m in range(len(members)):
        compare_member(ens, m, *want[m], "mhd")
        got = ens.download_global(m, xs)
        for n in xs:
            assert np.array_equal(got[n], want_b[m][n]), (m, n)
    ens.close()


def test_ensemble_workload_64_members_200_steps():
    """The 64-member workload (sizes 48..96 x 96..192, Re, theta, seeds drawn
    by ensemble_members), 200 steps: every member bit-identical to its own
    plain run."""
    import time
    members = pkg.ensemble_members(64)
    say("[gpu] 64-member ensemble, 200 steps")
    t0 = time.perf_counter()
    ens = pkg.Ensemble([pkg.Run(a) for a in members])
    say(f"[gpu]   set up in {time.perf_counter() - t0:.1f} s")
    ms = ens.autotune(2)
    say(f"[gpu]   autotuned {ens.tile_shape()} {ms:.3f} ms/step, {time.perf_counter() - t0:.1f} s")
    for done in range(0, 200, 50):
        ens.step(50)
        ens.sync()
        say(f"[gpu]   {done + 50} steps, {time.perf_counter() - t0:.1f} s")
    for m, a in enumerate(members):
        compare_member(ens, m, *single_run(a, 200), "64-member workload")
        if m % 8 == 7:
            say(f"[gpu]   members 0..{m} identical, {time.perf_counter() - t0:.1f} s")
    # the same 200 steps in one persistent launch, at the autotuned tiles
    pers = pkg.Ensemble([pkg.Run(a) for a in members])
    pers.set_tiles(*ens.tile_shape()[:3])
    pers.set_persistent(True)
    pers.step(200)
    pers.sync()
    say(f"[gpu]   persistent launch of 200 steps done, {time.perf_counter() - t0:.1f} s")
    for m in range(len(members)):
        want = ens.download_global(m, XZ_FIELDS)
        compare_member(pers, m, want, ens.scalars(m), "64-member workload, persistent")
    pers.close()
    ens.close()


class _MarginRun:
    """A Run whose configuration carries a time-step screen margin."""

    def __init__(self, args, margin):
        self._run = pkg.Run(args)
        self._margin = margin

    def cfg(self):
        c = self._run.cfg()
        c.dt_screen = self._margin
        return c

    def __getattr__(self, name):
        return getattr(self._run, name)


@pytest.mark.parametrize("tiles", [(8, 0), (4, 16), (9, 32), (16, 32, 32), (8, 0, 32),
                                   (28, 32, 16)],
                         ids=["ty8", "ty4_kc16", "ty9_kc32", "tw32_ty16", "tw32_ty8", "tw16_ty28"])
def test_persistent_launch_members_match_single_runs(tiles):
    """One cooperative launch per step() call: bit-identical to the members'
    own runs, also when alternated with graph replay."""
    say(f"[gpu] persistent ensemble of {len(SMALL)} uneven members, tiles {tiles}")
    want = [single_run(a, 7) for a in SMALL]
    ens = pkg.Ensemble([pkg.Run(a) for a in SMALL])
    ens.set_tiles(*tiles)
    ens.set_persistent(True)
    ens.step(3)
    ens.set_persistent(False)
    ens.step(1)
    ens.set_persistent(True)
    ens.step(2)
    ens.step(1)
    for m in range(len(SMALL)):
        compare_member(ens, m, *want[m], f"persistent, tiles {tiles}")
    ens.close()


def test_persistent_launch_exact_dt_pass():
    """Members whose screen margin sends every step through the exact dt pass
    (-0.5) next to members with the default margin, in one persistent launch:
    the redo phase runs for the flagged members only and every member equals
    its own run."""
    margins = [-0.5, 0.0, -0.5, 0.0]
    say("[gpu] persistent ensemble, exact dt pass on half the members")
    want = [single_run(a, 5) for a in SMALL]
    ens = pkg.Ensemble([_MarginRun(a, mg) for a, mg in zip(SMALL, margins)])
    ens.set_persistent(True)
    ens.step(5)
    for m in range(len(SMALL)):
        compare_member(ens, m, *want[m], f"persistent, margin {margins[m]}")
    ens.close()


def test_persistent_launch_mhd_members():
    mag = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2",
           "--bzp", "10", "--lrot", "1", "--ry", "0.1", "--ang", "0"]
    members = [
        ["--npx", "32", "--npy", "32", "--npz", "48", "--npey", "2", "--npez", "2", "--seed",
         "-11"] + mag,
        ["--npx", "70", "--npy", "24", "--npz", "40", "--npey", "1", "--npez", "2", "--seed",
         "-12", "--re", "3.5"] + mag,
    ]
    say("[gpu] persistent MHD ensemble")
    xs = ["bx", "by", "bz"]
    want, want_b = [], []
    for a in members:
        want.append(single_run(a, 4))
        sim = pkg.Run(a).make_sim()
        sim.step(4)
        want_b.append(sim.download_global(xs))
        sim.close()
    ens = pkg.Ensemble([pkg.Run(a) for a in members])
    ens.set_persistent(True)
    ens.step(4)
    for m in range(len(members)):
        compare_member(ens, m, *want[m], "persistent mhd")
        got = ens.download_global(m, xs)
        for n in xs:
            assert np.array_equal(got[n], want_b[m][n]), (m, n)
    ens.close()
