"""HIP time step against the oracle and the reference fixtures (bit-exact).

Each case builds the run with the product's host setup, steps it on the GPU
through the C ABI, and compares every virtual rank's downloaded arrays with
the oracle's (pinned to the reference CPU build): primaries over the whole
rank array including all ghost cells, RK accumulators over the interior,
and the scalars dt, timc, timt, umach, nit.  Where the reference fixture
holds the same step, the downloaded arrays are also hashed against it.
"""
import hashlib
import sys

import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

X_FIELDS = ["ru", "rv", "rw", "ro", "tt"]
Z_FIELDS = ["zru", "zrv", "zrw", "zro", "ztt"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:20]


def compare_with_oracle(sim, orc, tag):
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
    for k in ("dt", "timc", "timt", "umach"):
        if so[k] != sg[k]:
            bad.append((k, sg[k], so[k]))
    if so["nit"] != sg["nit"]:
        bad.append(("nit", sg["nit"], so["nit"]))
    assert not bad, f"{tag}: {bad[:6]}"


def check_golden(sim, rec, npey):
    """Hashes of the downloaded rank arrays against the reference dump."""
    bad = []
    for r, want in enumerate(rec):
        got = sim.download_rank(r, X_FIELDS + Z_FIELDS)
        for name in X_FIELDS + Z_FIELDS:
            if digest(got[name]) != want["fields"][name]:
                bad.append((r, name))
    sc = sim.scalars()
    for k in ("dt", "timc", "timt", "umach"):
        if sc[k].hex() != rec[0]["scalars"][k]:
            bad.append((k, sc[k]))
    return bad


GOLDEN_CASES = ["g16_2x2", "d32_1x2", "d32_4x2", "d32_2x4", "b_itcon0", "b_tp0",
                "b_itcon0_tp0", "b_itcon4", "b_izcon1", "b_rot", "b_tc", "b_ngrid0",
                "b_ngrid2", "b_pzp", "b_re_theta", "u40x24_2x2"]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_hip_matches_reference_fixtures(golden, name):
    cfg = golden[name]
    args = cfg["args"] + ["--npey", str(cfg["npey"]), "--npez", str(cfg["npez"])]
    say(f"[gpu] golden case {name} args={' '.join(args)}")
    run = pkg.Run(args)
    sim = run.make_sim()
    done = 0
    for st in sorted(int(s) for s in cfg["dumps"]):
        say(f"[gpu]   {name}: stepping to {st}")
        sim.step(st - done)
        done = st
        bad = check_golden(sim, cfg["dumps"][str(st)], cfg["npey"])
        assert not bad, f"{name} step {st}: {bad[:8]}"
    sim.close()


def test_hip_h64_1x2_100_steps(golden):
    """The CPU-sized headline case (hydro 64x64x128 on 1x2), 100 steps vs the reference."""
    cfg = golden["h64_1x2"]
    args = cfg["args"] + ["--npey", "1", "--npez", "2"]
    say("[gpu] h64_1x2 100 steps")
    run = pkg.Run(args)
    sim = run.make_sim()
    done = 0
    for st in (1, 2, 100):
        sim.step(st - done)
        done = st
        bad = check_golden(sim, cfg["dumps"][str(st)], 1)
        assert not bad, f"h64_1x2 step {st}: {bad[:8]}"


@pytest.mark.parametrize("nstages", [1, 2, 3])
def test_hip_truncated_steps_match_oracle(nstages):
    """step(nstages) leaves X and Z exactly as the reference's step(nstages)."""
    args = ["--npx", "24", "--npy", "16", "--npz", "40", "--npey", "2", "--npez", "2",
            "--lrot", "1", "--ry", "0.3", "--ang", "0.2"]
    say(f"[gpu] truncated step nstages={nstages}")
    run = pkg.Run(args)
    sim = run.make_sim()
    orc = oracle.Sim(args)
    sim.step(2)
    orc.step(2)
    compare_with_oracle(sim, orc, "after 2 full steps")
    sim.step(1, nstages)
    orc.step(1, nstages)
    compare_with_oracle(sim, orc, f"after a {nstages}-stage step")
    sim.step(3)
    orc.step(3)
    compare_with_oracle(sim, orc, "3 more full steps")


@pytest.mark.parametrize("args", [
    ["--npx", "70", "--npy", "18", "--npz", "30", "--npey", "3", "--npez", "3"],
    ["--npx", "130", "--npy", "9", "--npz", "12", "--npey", "1", "--npez", "2", "--izcon", "1"],
    ["--npx", "5", "--npy", "7", "--npz", "9", "--npey", "1", "--npez", "3", "--itcon", "0"],
], ids=["ragged_x", "two_x_tiles", "tiny"])
def test_hip_ragged_tiles_match_oracle(args):
    """Grids that leave partial x/y tiles and short z chunks."""
    say(f"[gpu] ragged case {' '.join(args)}")
    run = pkg.Run(args)
    sim = run.make_sim()
    orc = oracle.Sim(args)
    for n in (1, 4):
        sim.step(n)
        orc.step(n)
        compare_with_oracle(sim, orc, f"{args} after +{n}")


def test_upload_rank_round_trip():
    """Per-rank upload of the oracle's arrays then download gives them back."""
    args = ["--npx", "24", "--npy", "24", "--npz", "48", "--npey", "2", "--npez", "3"]
    say("[gpu] rank upload round trip")
    run = pkg.Run(args)
    orc = oracle.Sim(args)
    orc.step(2)
    cfg = run.cfg()
    sim = pkg.HipSim(cfg)
    sim.set_metrics(run.metrics())
    for r in range(orc.nranks):
        sim.upload_rank(r, {n: orc.field(r, n) for n in X_FIELDS})
    sc = orc.scalars()
    sim.set_scalars(sc["dt"], sc["timc"], sc["timt"], sc["timi"], sc["umach"], sc["nit"])
    for r in range(orc.nranks):
        got = sim.download_rank(r, X_FIELDS)
        for n in X_FIELDS:
            assert np.array_equal(got[n], orc.field(r, n)), (r, n)
    sim.step(2)
    orc.step(2)
    compare_with_oracle(sim, orc, "rank-uploaded state after 2 steps")


def test_decomposition_invariance_of_device_state():
    """One global device state serves any virtual decomposition."""
    base = ["--npx", "32", "--npy", "32", "--npz", "64"]
    outs = []
    for npey, npez in ((1, 2), (2, 2), (4, 4)):
        say(f"[gpu] decomposition {npey}x{npez}")
        run = pkg.Run(base + ["--npey", str(npey), "--npez", str(npez)])
        sim = run.make_sim()
        sim.step(3)
        outs.append(sim.download_global(X_FIELDS))
    for o in outs[1:]:
        for n in X_FIELDS:
            assert np.array_equal(o[n], outs[0][n])


def noisy_start(args, seed=5, amp=0.5):
    """The oracle's set-up state plus grid-scale velocity noise (amp * rho *
    U(-0.5, 0.5) on each momentum component), which viscosity damps by ~1-2%
    of max(ogamma*UU/T) per step; ghosts refreshed."""
    orc = oracle.Sim(args)
    rng = np.random.default_rng(seed)
    for r in range(orc.nranks):
        ro = orc.field(r, "ro")[1:-1, 1:-1, 1:-1]
        for n in ("ru", "rv", "rw"):
            f = orc.field(r, n)[1:-1, 1:-1, 1:-1]
            f += ro * amp * (rng.random(f.shape) - 0.5)
    orc.communicate()
    return orc


@pytest.mark.parametrize("margin", [0.0, 0.01, -0.5, 1e9],
                         ids=["default", "tight", "exact_pass_every_step", "every_cell"])
def test_dt_screen_matches_oracle(margin):
    """The last stage's time-step screen, whatever its margin, gives the
    oracle's dt and umach of every step bit for bit: the default margin, a
    tight one (both outcomes occur: the oracle's umach^2 falls by more than
    1% on some steps and by less on others), thresholds inside the previous
    extrema (every step takes the exact pass) and one that evaluates every
    cell."""
    args = ["--npx", "32", "--npy", "24", "--npz", "40", "--npey", "2", "--npez", "2"]
    say(f"[gpu] dt screen margin={margin}")
    orc = noisy_start(args)
    run = pkg.Run(args)
    sim = run.make_sim(dt_screen=margin)
    for r in range(orc.nranks):
        sim.upload_rank(r, {n: orc.field(r, n) for n in X_FIELDS})
    sc = orc.scalars()
    sim.set_scalars(sc["dt"], sc["timc"], sc["timt"], sc["timi"], sc["umach"], sc["nit"])
    um2 = []
    for s in range(8):
        sim.step(1)
        orc.step(1)
        compare_with_oracle(sim, orc, f"margin {margin}, step {s + 1}")
        um2.append(orc.scalars()["umach"] ** 2)
    ratios = [b / a for a, b in zip(um2[1:], um2[2:])]
    assert min(ratios) < 0.99 < max(ratios), ratios
    sim.close()
