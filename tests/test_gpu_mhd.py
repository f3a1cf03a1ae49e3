"""HIP MHD (LMAG) time step against the oracle, bit for bit.

The reference has no runnable LMAG branch, so the oracle's restatement of the
Fortran FLUXES (oracle/r3d_oracle_mag.inc, pinned by the invariants of
test_oracle_mhd.py) is the comparison point: every primary over the whole rank
array including ghosts, the RK accumulators over the interior and the
scalars must be identical after each checked step.
"""
import sys

import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

X_FIELDS = ["ru", "rv", "rw", "ro", "tt", "bx", "by", "bz"]
Z_FIELDS = ["z" + n for n in X_FIELDS]
MAG = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2", "--bzp", "10"]
ROT = ["--lrot", "1", "--ry", "0.1", "--ang", "0"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare(sim, orc, tag):
    bad = []
    for r in range(orc.nranks):
        got = sim.download_rank(r, X_FIELDS + Z_FIELDS)
        for name in X_FIELDS + Z_FIELDS:
            want = orc.field(r, name)
            g = got[name]
            if name.startswith("z"):
                want = want[1:-1, 1:-1, 1:-1]
                g = g[1:-1, 1:-1, 1:-1]
            if not np.array_equal(g.view(np.uint64), want.view(np.uint64)):
                diff = np.abs(g - want)
                bad.append((r, name, int((g != want).sum()), float(diff.max()),
                            np.unravel_index(int(np.argmax(diff)), diff.shape)))
    so, sg = orc.scalars(), sim.scalars()
    for k in ("dt", "timc", "timt", "umach", "nit"):
        if so[k] != sg[k]:
            bad.append((k, sg[k], so[k]))
    assert not bad, f"{tag}: {bad[:6]}"


def randomize_field(orc, seed):
    """Give BX, BY, BZ smooth periodic structure plus noise on every rank
    (interior), then refresh the oracle's halos."""
    rng = np.random.default_rng(seed)
    nry, nrz = orc.npy // orc.npey, orc.npz // orc.npez
    for r in range(orc.nranks):
        py, pz = orc.rank_coords(r)
        k = (pz * nrz + np.arange(1, nrz + 1))[:, None, None]
        j = (py * nry + np.arange(1, nry + 1))[None, :, None]
        i = np.arange(1, orc.npx + 1)[None, None, :]
        x, y = 2 * np.pi * i / orc.npx, 2 * np.pi * j / orc.npy
        z = np.pi * k / orc.npz
        fields = {
            "bx": 0.4 * np.sin(x + 2 * y) * np.cos(z) + 0.1,
            "by": 0.3 * np.cos(2 * x - y) * np.sin(2 * z),
            "bz": 0.2 * np.sin(x) * np.sin(y) * np.sin(3 * z),
        }
        for name, v in fields.items():
            noise = rng.uniform(-1e-3, 1e-3, (nrz, nry, orc.npx))
            orc.field(r, name)[1:-1, 1:-1, 1:-1] = v + noise
    orc.communicate()


def upload_oracle_state(run, orc, tile_rows=0):
    """tile_rows: 0 = the library's MHD tile (12 rows x 32 columns), 5 = the
    5-row, 64-column one."""
    cfg = run.cfg()
    cfg.tile_rows = tile_rows
    sim = pkg.HipSim(cfg)
    sim.set_metrics(run.metrics())
    for r in range(orc.nranks):
        sim.upload_rank(r, {n: orc.field(r, n) for n in X_FIELDS})
    sc = orc.scalars()
    sim.set_scalars(sc["dt"], sc["timc"], sc["timt"], sc["timi"], sc["umach"], sc["nit"])
    return sim


def test_mhd_magnetic_layer_matches_oracle():
    """MHD + rotation physics of the 252x252x512 configuration on a small grid."""
    args = ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "4"] + MAG + ROT
    say(f"[gpu] mhd layer {' '.join(args)}")
    run = pkg.Run(args)
    sim = run.make_sim()
    orc = oracle.Sim(args)
    for n in (1, 1, 3):
        sim.step(n)
        orc.step(n)
        compare(sim, orc, f"layer after +{n}")


@pytest.mark.parametrize("tile_rows", [0, 5], ids=["tile12x32", "tile5x64"])
@pytest.mark.parametrize("nstages", [1, 2, 3])
def test_mhd_random_field_matches_oracle(nstages, tile_rows):
    """A fully three-dimensional B exercises every induction and Lorentz term."""
    args = ["--npx", "24", "--npy", "16", "--npz", "40", "--npey", "2", "--npez", "2"] + MAG + ROT
    say(f"[gpu] mhd random field nstages={nstages} tile_rows={tile_rows}")
    run = pkg.Run(args)
    orc = oracle.Sim(args)
    randomize_field(orc, 7)
    sim = upload_oracle_state(run, orc, tile_rows)
    compare(sim, orc, "after upload")
    sim.step(1, nstages)
    orc.step(1, nstages)
    compare(sim, orc, f"after a {nstages}-stage step")
    sim.step(2)
    orc.step(2)
    compare(sim, orc, "2 more full steps")


@pytest.mark.parametrize("args", [
    ["--npx", "16", "--npy", "16", "--npz", "24", "--npey", "2", "--npez", "1"],
    ["--npx", "70", "--npy", "18", "--npz", "30", "--npey", "3", "--npez", "3"],
    ["--npx", "9", "--npy", "8", "--npz", "9", "--npey", "2", "--npez", "3", "--izcon", "1"],
], ids=["single_z_rank", "ragged_x", "three_planes"])
@pytest.mark.parametrize("tile_rows", [0, 5], ids=["tile12x32", "tile5x64"])
def test_mhd_edge_geometries_match_oracle(args, tile_rows):
    """One z rank (no one-sided bottom stencils), partial tiles, 3-plane ranks."""
    args = args + MAG
    say(f"[gpu] mhd edge case {' '.join(args)} tile_rows={tile_rows}")
    run = pkg.Run(args)
    orc = oracle.Sim(args)
    randomize_field(orc, 11)
    sim = upload_oracle_state(run, orc, tile_rows)
    for n in (1, 2):
        sim.step(n)
        orc.step(n)
        compare(sim, orc, f"{args} after +{n}")


def test_mhd_zero_field_matches_oracle():
    """ampb = 0: the magnetic branch runs on B = 0."""
    args = ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2",
            "--lmag", "1", "--rm", "50", "--beta", "10", "--bfh", "2", "--bzp", "10"]
    say("[gpu] mhd zero field")
    run = pkg.Run(args)
    sim = run.make_sim()
    orc = oracle.Sim(args)
    sim.step(2)
    orc.step(2)
    compare(sim, orc, "B = 0 after 2 steps")
    got = sim.download_global(["bx", "by", "bz"])
    assert not any(got[n].any() for n in got)
