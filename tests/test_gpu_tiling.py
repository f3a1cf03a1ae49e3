"""The stage kernel at every built tiling against the reference fixtures and the oracle.

The library picks its stage tiling from the grid (64-lane x tiles, 8 or 9
rows per hydro tile, 5 with LMAG, z chunks of 8..128 planes so that a launch
has enough workgroups).  Small test grids would all get 8-plane chunks, so
here the tiling is forced (r3d_hip_cfg.tile_rows / zchunk) and each forced
shape must reproduce the reference's dumps (tests/golden/ref_hashes.json) or
the oracle bit for bit: the z-chunk staging of the metrics and the extra
planes at full depth (128), the chunk seams, partial tiles in x, y and z.
"""
import hashlib
import sys

import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

X5 = ["ru", "rv", "rw", "ro", "tt"]
Z5 = ["zru", "zrv", "zrw", "zro", "ztt"]
MAG = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2", "--bzp", "10",
       "--lrot", "1", "--ry", "0.1", "--ang", "0"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:20]


def check_golden(sim, rec):
    bad = []
    for r, want in enumerate(rec):
        got = sim.download_rank(r, X5 + Z5)
        for name in X5 + Z5:
            if digest(got[name]) != want["fields"][name]:
                bad.append((r, name))
    sc = sim.scalars()
    for k in ("dt", "timc", "timt", "umach"):
        if sc[k].hex() != rec[0]["scalars"][k]:
            bad.append((k, sc[k]))
    return bad


TILINGS = [(8, 16), (9, 32), (8, 64), (9, 128), (8, 128)]


@pytest.mark.parametrize("tiling", TILINGS, ids=[f"ty{t}_kc{k}" for t, k in TILINGS])
@pytest.mark.parametrize("name", ["h64_1x2", "g16_2x2", "d32_4x2", "b_rot", "b_izcon1",
                                  "u40x24_2x2", "b_ngrid2"])
def test_forced_tiling_matches_reference(golden, name, tiling):
    ty, kc = tiling
    cfg = golden[name]
    args = cfg["args"] + ["--npey", str(cfg["npey"]), "--npez", str(cfg["npez"])]
    say(f"[gpu] {name} ty={ty} kc={kc}")
    run = pkg.Run(args)
    sim = run.make_sim(tile_rows=ty, zchunk=kc)
    done = 0
    for st in sorted(int(s) for s in cfg["dumps"] if int(s) > 0):
        sim.step(st - done)
        done = st
        bad = check_golden(sim, cfg["dumps"][str(st)])
        assert not bad, f"{name} ty={ty} kc={kc} step {st}: {bad[:8]}"
    sim.close()


def compare_oracle(sim, orc, names, tag):
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


@pytest.mark.parametrize("kc", [16, 32, 64, 128])
def test_mhd_forced_chunks_match_oracle(kc):
    """Rotating MHD on a grid whose z extent (136) is not a multiple of the chunk."""
    args = ["--npx", "40", "--npy", "20", "--npz", "136", "--npey", "2", "--npez", "4"] + MAG
    say(f"[gpu] MHD kc={kc}")
    run = pkg.Run(args)
    sim = run.make_sim(zchunk=kc)
    orc = oracle.Sim(args)
    sim.step(3)
    orc.step(3)
    names = X5 + ["bx", "by", "bz"]
    compare_oracle(sim, orc, names + ["z" + n for n in names], f"MHD kc={kc}")
    sim.close()


@pytest.mark.parametrize("kc", [16, 128])
@pytest.mark.parametrize("physics", ["heat", "density"])
def test_feature_kernels_forced_chunks_match_oracle(physics, kc):
    extra = {"heat": ["--itcon", "2", "--tp", "1.5", "--zp", "27.0", "--hh", "6.0"],
             "density": ["--ngrid", "0", "--id", "1", "--dh", "0.4", "--dv", "0.3"]}[physics]
    args = ["--npx", "40", "--npy", "30", "--npz", "140", "--npey", "2", "--npez", "2"] + extra
    say(f"[gpu] {physics} kc={kc}")
    run = pkg.Run(args)
    sim = run.make_sim(zchunk=kc)
    orc = oracle.Sim(args)
    sim.step(2)
    orc.step(2)
    compare_oracle(sim, orc, X5 + Z5, f"{physics} kc={kc}")
    sim.close()

