"""The CPU oracle against dumps of the real reference CPU build.

tests/golden/ref_hashes.json holds, per configuration and dump step, a hash
of every field (whole rank array with ghosts), every 1-D metric and the exact
scalars of the reference build (oracle/_ref/3dmhd, produced by
tests/golden/make_golden.py).  The oracle must reproduce all of them, so each
of its sweeps, the halo exchange, BCON branches, the dt reduction and the
initial condition are pinned bit-for-bit.
"""
import hashlib

import numpy as np
import pytest

import oracle

Y_METRICS = ("wyy", "dyydy", "d2yydy2", "ddy")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:20]


def compare(sim, want, npey):
    bad = []
    for r, rec in enumerate(want):
        for name, h in rec["fields"].items():
            if digest(sim.field(r, name)) != h:
                bad.append((r, name))
        for name, h in rec["metrics"].items():
            v = sim.metric(r, name)
            if npey == 1 and name in Y_METRICS:
                v = v[: sim.ny - 1]
            if digest(v) != h:
                bad.append((r, "metric:" + name))
        sc = sim.scalars()
        for k, hx in rec["scalars"].items():
            if sc[k].hex() != hx:
                bad.append((r, k, sc[k].hex(), hx))
        assert sc["nit"] == rec["nit"]
    return bad


CASES = ["g16_2x2", "d32_1x2", "d32_2x2", "d32_4x2", "d32_2x4", "d32_1x4",
         "b_itcon0", "b_tp0", "b_itcon0_tp0", "b_itcon4", "b_izcon1", "b_rot", "b_tc",
         "b_ngrid0", "b_ngrid2", "b_pzp", "b_re_theta", "u40x24_2x2", "h64_1x2"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_dumps(golden, name):
    cfg = golden[name]
    sim = oracle.Sim(cfg["args"] + ["--npey", str(cfg["npey"]), "--npez", str(cfg["npez"])])
    done = 0
    for st in sorted(int(s) for s in cfg["dumps"]):
        sim.step(st - done)
        done = st
        bad = compare(sim, cfg["dumps"][str(st)], cfg["npey"])
        assert not bad, f"{name} step {st}: {bad[:8]}"


def test_planted_mistake_is_caught(golden):
    """A one-ulp change to one interior value must break the field hash."""
    cfg = golden["g16_2x2"]
    sim = oracle.Sim(cfg["args"] + ["--npey", "2", "--npez", "2"])
    sim.step(1)
    tt = sim.field(0, "tt")
    tt[3, 3, 3] = np.nextafter(tt[3, 3, 3], np.inf)
    bad = compare(sim, cfg["dumps"]["1"], 2)
    assert (0, "tt") in bad
