"""Known-answer tests of the oracle's setup pieces.

Values come from the reference's own tests (tests/params_test.cpp:11-51,
84-142, tests/ran2_ode_test.cpp:49-72, tests/_ran2_ref.inc,
tests/grid_test.cpp:122-140) or from closed forms.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def f32(x):
    return float(np.float32(x))


def test_ran2_golden_draws():
    ref = json.load(open(os.path.join(HERE, "golden", "ran2_ref.json")))["draws"]
    seq = oracle.ran2_sequence(30000)
    for n, v in ref:
        assert seq[n - 1] == v, n


def test_ran2_is_uniform_in_unit_interval():
    seq = oracle.ran2_sequence(200000)
    assert seq.min() >= 0.0 and seq.max() < 1.0  # iiy*RM with iiy in [0, M)
    assert abs(seq.mean() - 0.5) < 0.005


def test_static_profile_golden():
    s = oracle.Sim(["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2"], init=False)
    t, r = s.static_profile()
    assert t[0] == 1.0
    assert t[3] == 3.21832888589760424
    assert t[9] == 6.00075810678302801
    assert t[19] == 8.71388943224966717
    assert t[31] == 10.9999998343435870
    assert r[31] == 36.4828718697596557
    assert np.all(np.diff(t) > 0) and np.all(np.diff(r) > 0)


def test_derived_golden_geometry():
    s = oracle.Sim(["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2"], init=False)
    assert (s.nx, s.ny, s.nz) == (18, 10, 18)
    d = s.derived()
    assert d["pr"] == 0.05000000074505806
    assert d["ampt"] == 0.004999999888241291
    assert d["sf"] == 0.31499999761581421
    assert d["gamma"] == 1.6666666269302368
    assert d["c13"] == f32(1.0 / 3.0) and d["c23"] == 2.0 * f32(1.0 / 3.0)
    assert d["c43"] == 4.0 * f32(1.0 / 3.0)
    assert d["gam1"] == f32(np.float32(8.0) / np.float32(15.0))
    assert d["gam2"] == f32(np.float32(5.0) / np.float32(12.0))
    assert d["gam3"] == 0.75
    assert d["zeta1"] == f32(np.float32(-17.0) / np.float32(60.0))
    assert d["zeta2"] == f32(np.float32(-5.0) / np.float32(12.0))
    assert (d["hx"], d["h2x"], d["hz"], d["h2z"]) == (7.5, 225.0, 15.5, 961.0)
    assert d["repr"] == 2.0 * f32(0.05) and d["ore"] == 0.5
    assert d["cv"] == 1.0 / (1.6666666269302368 - 1.0)
    assert d["rkapst"] == 0.25 / 0.625
    assert (d["orm"], d["obeta"], d["omx"], d["omz"]) == (0.0, 0.0, 0.0, 0.0)


def test_derived_magnetized_rotating():
    s = oracle.Sim(["--lmag", "1", "--rm", "4", "--beta", "2", "--lrot", "1", "--ry", "2",
                    "--ang", "0.5", "--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2"],
                   init=False)
    d = s.derived()
    assert (d["orm"], d["obeta"]) == (0.25, 1.0)
    assert d["omx"] == math.sin(0.5) / 2.0 and d["omz"] == math.cos(0.5) / 2.0


def test_mkgrid_closed_forms():
    u = oracle.Sim(["--ngrid", "0", "--npx", "16", "--npy", "16", "--npz", "32"], init=False)
    assert u.mkgrid(0.5, 1) == (10.0, 1.0 / 20.0, 0.0)
    assert u.mkgrid(0.0, 2)[0] == 0.0
    assert u.mkgrid(1.0, 3)[0] == 40.0
    a = oracle.Sim(["--npx", "16", "--npy", "16", "--npz", "32"], init=False)
    s, d1, d2 = a.mkgrid(1.0 / 15.0, 1)
    assert abs(s - 5.8619437) < 1e-5
    # arctan map: x(s) = (tan(a2 s + a3) - s1)/a1 and its derivative in closed form
    s1, s2, smax = -7.0, 7.0, 20.0
    a1 = (s2 - s1) / smax
    a3 = math.atan(s1)
    a2 = math.atan(s2) - a3
    for sc in (0.1, 0.37, 0.8):
        x, dx, _ = a.mkgrid(sc, 1)
        assert x == pytest.approx((math.tan(a2 * sc + a3) - s1) / a1, rel=1e-14)
        assert dx == pytest.approx(a1 / a2 / (1 + (s1 + a1 * x) ** 2), rel=1e-14)


def test_unknown_flag_rejected():
    with pytest.raises(oracle.OracleError, match="unknown option"):
        oracle.Sim(["--bogus", "1"], init=False)
