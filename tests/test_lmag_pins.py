"""The oracle's LMAG set-up and time-step limits against the reference's own code.

The reference C++ cannot take an LMAG step: its FLUXES stops
(src/fluxes.cpp:539-541).  Two parts of the LMAG path run before that stop
and are pinned here against the reference build (tests/golden/ref_lmag.json,
written by tests/golden/make_golden.py from oracle/_ref):

* the initial state with the magnetic layer and its density correction
  (src/static_ic.cpp:94-108): the reference writes gstate.start
  (src/main.cpp:199) before its first step, so every field, metric and
  scalar of that dump is hashed;
* STEP's time-step limits (src/step.cpp:121-229): UU with the fast
  magnetosonic speed, the resistive limit VV = 0.5 dd^2 Rm, WW1..WW3, umach
  and dt = sff * min of the four.  oracle/_ref/stepdt runs the reference's
  own STEP and captures them at its first FLUXES call, at rest and with a
  deterministic flow and transverse field (oracle/ref_stepdt.cpp).

Closed forms pin the same limits independently of the reference: for a
uniform state the fast speed is (|u| + sqrt(gamma T + b2))^2 with
b2 = |B|^2 obeta / rho, so dt is sff times the least of dd/(|u| + c_f),
0.5 dd^2 repr cv rho / kappa, 0.375 dd^2 Re rho and 0.5 dd^2 Rm.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
STEPDT_FIELDS = ["ru", "rv", "rw", "ro", "tt", "bx", "by", "bz", "uu", "vv", "ww1", "ww2", "ww3"]
Y_METRICS = ("wyy", "dyydy", "d2yydy2", "ddy")


@pytest.fixture(scope="module")
def ref():
    with open(os.path.join(HERE, "golden", "ref_lmag.json")) as f:
        return json.load(f)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:20]


def apply_flow(sim):
    """The deterministic flow of oracle/ref_stepdt.cpp on every rank's interior."""
    nry, nrz = sim.npy // sim.npey, sim.npz // sim.npez
    for r in range(sim.nranks):
        py, pz = sim.rank_coords(r)
        k = np.arange(1, sim.nz - 1, dtype=np.int64)[:, None, None]
        j = np.arange(1, sim.ny - 1, dtype=np.int64)[None, :, None]
        i = np.arange(1, sim.nx - 1, dtype=np.int64)[None, None, :]
        jg, kg = py * nry + j, pz * nrz + k
        w = [((i * 7919 + jg * 104729 + kg * 1299709 + c * 15485863) % 10007) / 10007.0 - 0.5
             for c in range(5)]
        ro = sim.field(r, "ro")[1:-1, 1:-1, 1:-1]
        for c, name in enumerate(("ru", "rv", "rw")):
            sim.field(r, name)[1:-1, 1:-1, 1:-1] = ro * (w[c] * 0.4)
        if sim.lmag:
            sim.field(r, "by")[1:-1, 1:-1, 1:-1] = w[3] * 4.0
            sim.field(r, "bz")[1:-1, 1:-1, 1:-1] = w[4] * 4.0


@pytest.mark.parametrize("name", ["m16_2x4", "m252_2x4"])
def test_oracle_lmag_start_state_matches_reference_dump(ref, name):
    cfg = ref["start"][name]
    sim = oracle.Sim(cfg["args"] + ["--npey", str(cfg["npey"]), "--npez", str(cfg["npez"])])
    bad = []
    for r, rec in enumerate(cfg["dumps"]["0"]):
        for f, h in rec["fields"].items():
            if digest(sim.field(r, f)) != h:
                bad.append((r, f))
        for m, h in rec["metrics"].items():
            v = sim.metric(r, m)
            if cfg["npey"] == 1 and m in Y_METRICS:
                v = v[: sim.ny - 1]
            if digest(v) != h:
                bad.append((r, "metric:" + m))
    sc = sim.scalars()
    for k, hx in cfg["dumps"]["0"][0]["scalars"].items():
        if sc[k].hex() != hx:
            bad.append((k, sc[k].hex(), hx))
    assert not bad, bad[:8]
    # the layer is there: B is at most ampb = 1, peaks near z = bzp and falls
    # off away from it
    bx = sim.global_interior("bx")
    assert 0.3 < bx.max() <= 1.0 and bx.min() < 0.1 * bx.max()


def test_lmag_start_fixture_would_catch_a_one_ulp_change(ref):
    cfg = ref["start"]["m16_2x4"]
    sim = oracle.Sim(cfg["args"] + ["--npey", "2", "--npez", "4"])
    bx = sim.field(5, "bx")
    bx[4, 4, 4] = np.nextafter(bx[4, 4, 4], np.inf)
    assert digest(sim.field(5, "bx")) != cfg["dumps"]["0"][5]["fields"]["bx"]


STEPDT_CPU = ["sdt_m16_rest", "sdt_m16_flow", "sdt_m16_rm", "sdt_m16_beta", "sdt_m16_adv_rest",
              "sdt_m16_adv_flow", "sdt_m16_adv_beta", "sdt_m24_flow_1x2", "sdt_h24_flow",
              "sdt_m252_flow"]


@pytest.mark.parametrize("name", STEPDT_CPU)
def test_oracle_step_limits_match_reference(ref, name):
    cfg = ref["stepdt"][name]
    sim = oracle.Sim(cfg["args"] + ["--npey", str(cfg["npey"]), "--npez", str(cfg["npez"])])
    if cfg["flow"]:
        apply_flow(sim)
    sim.step_dt()
    sc = sim.scalars()
    bad = []
    for r, rec in enumerate(cfg["ranks"]):
        for f in STEPDT_FIELDS:
            if f == "vv" and not sim.lmag:
                continue  # not a limit without LMAG
            if digest(sim.field(r, f)[1:-1, 1:-1, 1:-1]) != rec["interior"][f]:
                bad.append((r, f))
        if sc["dt"].hex() != rec["dt"] or sc["umach"].hex() != rec["umach"]:
            bad.append((r, "dt/umach", sc["dt"].hex(), rec["dt"]))
    assert not bad, bad[:8]


def test_advective_cases_bind_on_the_fast_speed(ref):
    """The dt of the Re*Pr = 50 cases depends on the field and the flow, so a
    GPU that matches them uses b2 and the flow term in its dt."""
    dts = {n: ref["stepdt"][n]["ranks"][0]["dt"]
           for n in ("sdt_m16_adv_rest", "sdt_m16_adv_flow", "sdt_m16_adv_beta")}
    assert len(set(dts.values())) == 3, dts
    assert ref["stepdt"]["sdt_m16_rm"]["ranks"][0]["dt"] != dts["sdt_m16_adv_rest"]


# ---- closed forms -----------------------------------------------------------

UNIFORM = ["--ngrid", "0", "--npx", "12", "--npy", "12", "--npz", "24", "--npey", "2",
           "--npez", "2", "--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "0",
           "--ampt", "0"]


def uniform_state(sim, rho, T, u, B):
    for r in range(sim.nranks):
        sl = (slice(1, -1),) * 3
        sim.field(r, "ro")[sl] = rho
        sim.field(r, "tt")[sl] = T
        for c, name in enumerate(("ru", "rv", "rw")):
            sim.field(r, name)[sl] = rho * u[c]
        for c, name in enumerate(("bx", "by", "bz")):
            sim.field(r, name)[sl] = B[c]


def closed_form(sim, rho, T, u, B, mistake=None):
    """dt and the fast speed of a uniform state (kappa = 1 with pzp = 0)."""
    d = sim.derived()
    gamma, obeta = d["gamma"], d["obeta"]
    re, rm = 1.0 / d["ore"], 1.0 / d["orm"]
    dd = min(sim.metric(0, "ddx")[1], sim.metric(0, "ddy")[1], sim.metric(0, "ddz")[1])
    speed = float(np.sqrt(np.dot(u, u)))
    b2 = float(np.dot(B, B)) * obeta / rho
    if mistake == "b2_without_obeta":
        b2 = float(np.dot(B, B)) / rho
    if mistake == "b2_times_rho":
        b2 = float(np.dot(B, B)) * obeta * rho
    cf = np.sqrt(gamma * T + b2)
    fast = (speed + cf) ** 2
    if mistake == "no_cross_term":
        fast = speed ** 2 + cf ** 2
    if mistake == "b2_outside_sqrt":
        fast = (speed + np.sqrt(gamma * T)) ** 2 + b2
    lim = [dd / np.sqrt(fast), 0.5 * dd * dd * d["repr"] * d["cv"] * rho,
           0.375 * dd * dd * re * rho, 0.5 * dd * dd * (re if mistake == "vv_with_re" else rm)]
    return d["sf"] * min(lim), fast, int(np.argmin(lim))


# (extra flags, rho, T, u, B, limit index that binds)
CASES = [
    (["--re", "50", "--pr", "1"], 1.7, 2.3, (0.3, -0.2, 0.1), (1.5, -0.7, 0.4), 0),
    (["--re", "50", "--pr", "1"], 0.8, 1.1, (0.0, 0.0, 0.0), (2.5, 1.0, -1.0), 0),
    ([], 1.7, 2.3, (0.3, -0.2, 0.1), (1.5, -0.7, 0.4), 1),
    (["--re", "0.05", "--pr", "1"], 1.7, 2.3, (0.3, -0.2, 0.1), (1.5, -0.7, 0.4), 2),
    (["--re", "50", "--pr", "1", "--rm", "0.02"], 1.7, 2.3, (0.3, -0.2, 0.1), (1.5, -0.7, 0.4), 3),
]


def oracle_dt(flags, rho, T, u, B):
    sim = oracle.Sim(UNIFORM + flags)
    uniform_state(sim, rho, T, u, B)
    sim.step_dt()
    return sim, sim.scalars()["dt"], sim.field(0, "uu")[1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_step_limits_closed_form(case):
    flags, rho, T, u, B, binds = CASES[case]
    sim, dt, uu = oracle_dt(flags, rho, T, u, B)
    want, fast, which = closed_form(sim, rho, T, u, B)
    assert which == binds
    assert abs(dt - want) <= 1e-14 * want, (dt, want)
    np.testing.assert_allclose(uu, fast, rtol=1e-14, atol=0)


@pytest.mark.parametrize("mistake", ["b2_without_obeta", "b2_times_rho", "no_cross_term",
                                     "b2_outside_sqrt", "vv_with_re"])
def test_closed_form_catches_planted_mistakes(mistake):
    """Each mistake changes dt in at least one case beyond the tolerance, so a
    limit computed with it could not pass the closed-form check."""
    caught = False
    for flags, rho, T, u, B, _ in CASES:
        sim, dt, _ = oracle_dt(flags, rho, T, u, B)
        wrong, _, _ = closed_form(sim, rho, T, u, B, mistake)
        caught |= abs(dt - wrong) > 1e-14 * abs(dt)
    assert caught, mistake
