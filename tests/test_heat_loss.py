"""ITCON=2, the localized embedded heat loss (CPU: oracle and host setup).

The Fortran code adds to the energy flux, over every interior cell,
  FT -= TP*OCV*[0.5*(1+TANH(LOG(3)/QFH*(TIMT-TC)))]*(1/RO)
        *EXP(CLN*(EXX-XP)**2)/HH*EXP(CLN*(WYY-YP)**2)/HH*EXP(CLN*(ZEE-ZP)**2)/HH
with CLN = -4*LOG(2)/HH**2 (golden/fortran/3dmhdsub.f:1797-1822; the bracket
only when TC != 0).  The reference C++ refuses ITCON=2 (src/fluxes.cpp:50),
so no reference output exists for it: the oracle's term is pinned here by
what the formula fixes, each check written in a form other than the
oracle's:

* on a uniform fluid at rest every other energy term vanishes, so FT is the
  heat loss alone: a Gaussian of full width at half maximum HH centred on
  (XP, YP, ZP), amplitude -TP/(CV*RO*HH**3), evaluated here as
  2**(-4*((x-XP)/HH)**2) per axis with global grid coordinates;
* FT scales exactly as 1/RO (doubling RO halves FT bit for bit);
* the ramp: TC = QFH at TIMT = 0 gives 0.5*(1+tanh(-ln 3)) = 0.1 of the
  unramped loss (tanh(ln 3) = 4/5);
* TP = 0 leaves the whole step bit-identical to ITCON=0 with TP=0, whose
  step is pinned by the reference fixture b_itcon0_tp0 (top boundary: both
  hold the temperature fixed, src/bcon.cpp:71);
* ITCON=3 (only named in the Fortran, 3dmhd.f:645) stays refused.
"""
import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()

BASE = ["--npx", "12", "--npy", "10", "--npz", "16", "--npey", "1", "--npez", "2"]
HEAT = ["--itcon", "2", "--tp", "1.5", "--xp", "10.2", "--yp", "9.9", "--zp", "27.0",
        "--hh", "6.0"]


def at_rest(sim, rho, temp=1.3):
    for r in range(sim.nranks):
        for n in ("ru", "rv", "rw"):
            sim.field(r, n)[...] = 0.0
        sim.field(r, "ro")[...] = rho
        sim.field(r, "tt")[...] = temp


def heat_ft(args, rho):
    """FT of every rank (interior cells) after fluxes on a uniform state."""
    sim = oracle.Sim(args)
    at_rest(sim, rho)
    out = []
    for r in range(sim.nranks):
        sim.fluxes(r)
        out.append(sim.field(r, "ft")[1:-1, 1:-1, 1:-1].copy())
    return sim, out


def value(args, name, default):
    return float(args[args.index(name) + 1]) if name in args else default


def expected(sim, r, args, rho, ramp=1.0):
    tp, hh = value(args, "--tp", 2.0), value(args, "--hh", 1.0)
    xp, yp, zp = value(args, "--xp", 10.0), value(args, "--yp", 10.0), value(args, "--zp", 0.0)
    cv = sim.derived()["cv"]
    x = sim.metric(r, "exx")[1:-1]
    y = sim.metric(r, "wyy")[1:-1]
    z = sim.metric(r, "zee")[1:-1]
    g = [2.0 ** (-4.0 * ((c - c0) / hh) ** 2) for c, c0 in ((x, xp), (y, yp), (z, zp))]
    return -(tp / cv) * ramp / rho / hh ** 3 * g[2][:, None, None] * g[1][None, :, None] * \
        g[0][None, None, :]


def test_heat_loss_is_the_gaussian_sink_on_a_fluid_at_rest():
    rho = 0.7
    sim, ft = heat_ft(BASE + HEAT, rho)
    for r in range(sim.nranks):
        want = expected(sim, r, BASE + HEAT, rho)
        # single-precision LOG(2) in CLN: relative error of the exponent ~ 3e-8
        np.testing.assert_allclose(ft[r], want, rtol=2e-6, atol=0)
        assert (ft[r] < 0).all()
    # the sink reaches both z ranks (ZP lies near their boundary)
    assert min(abs(f).max() for f in ft) > 1e-3 * max(abs(f).max() for f in ft)


def test_heat_loss_scales_as_one_over_density_exactly():
    _, a = heat_ft(BASE + HEAT, 0.7)
    _, b = heat_ft(BASE + HEAT, 1.4)
    for fa, fb in zip(a, b):
        assert np.array_equal(fa, 2.0 * fb)


def test_heat_loss_ramp():
    rho = 0.9
    _, full = heat_ft(BASE + HEAT, rho)
    sim, ramped = heat_ft(BASE + HEAT + ["--tc", "0.25", "--qfh", "0.25"], rho)
    assert sim.scalars()["timt"] == 0.0
    for f, g in zip(full, ramped):
        np.testing.assert_allclose(g, 0.1 * f, rtol=1e-6, atol=0)


def test_zero_amplitude_is_the_fixed_temperature_run():
    args0 = BASE + ["--itcon", "0", "--tp", "0"]
    args2 = BASE + ["--itcon", "2", "--tp", "0", "--zp", "27.0", "--hh", "6.0"]
    a, b = oracle.Sim(args0), oracle.Sim(args2)
    a.step(3)
    b.step(3)
    for r in range(a.nranks):
        for n in ("ru", "rv", "rw", "ro", "tt", "zru", "zrv", "zrw", "zro", "ztt"):
            assert np.array_equal(a.field(r, n).view(np.uint64), b.field(r, n).view(np.uint64)), n
    assert a.scalars() == b.scalars()


def test_heat_loss_changes_the_step_only_near_the_sink():
    a = oracle.Sim(BASE + ["--itcon", "0", "--tp", "0"])
    b = oracle.Sim(BASE + HEAT)
    a.step(1)
    b.step(1)
    ta, tb = a.global_interior("tt"), b.global_interior("tt")
    assert (tb[1:-1] <= ta[1:-1]).all()  # a sink only cools (boundary planes held)
    assert (tb[1:-1] < ta[1:-1]).any()


def test_host_accepts_itcon_2_and_refuses_itcon_3():
    run = pkg.Run(BASE + HEAT + ["--qfh", "0.5"])
    cfg = run.cfg()
    assert cfg.itcon == 2 and cfg.zp == 27.0 and cfg.qfh == 0.5
    with pytest.raises(pkg.HipError, match="ITCON=3"):
        pkg.Run(BASE + ["--itcon", "3"])
    with pytest.raises(pkg.HipError, match="QFH"):
        pkg.Run(BASE + HEAT + ["--tc", "0.5"])
    with pytest.raises(RuntimeError, match="ITCON=3"):
        oracle.Sim(BASE + ["--itcon", "3"]).step(1)
