"""LREM (M. Rempel's stratification) in the oracle, pinned by closed forms.

The reference C++ refuses LREM (src/static_ic.cpp:27, src/fluxes.cpp:47);
its only specification is the Fortran (golden/fortran/3dmhdsub.f:437-490
STATIC, 817-851 STEP, 1316 and 1534-1596 FLUXES, 2852-2952 BCON, 4442-4650
the background ODE; golden/fortran/3dmhdset.f:154-170 KAPPA), which cannot be
compiled here, so the oracle's restatement is parity unpinned against it and
is pinned by what the mathematics fixes:

* the background T(z), p = rho T solves dT/dz = 0.4/kappa or 0.41 (the
  smaller-gradient switch), dp/dz = p/T, T = p = 1 at z = 2, kappa =
  rho + (1.25 - rho) exp(-5 z): checked by differences of the solution, and
  in an adiabatic stretch by T linear with slope 0.41 and p = p0 (T/T0)^(1/0.41);
* RK, DRK, TU, DZTU, DZTB against the same derivatives;
* FLUXES: for planes of uniform temperature LREM's energy flux equals the
  ordinary conduction with RKAPA/DKAPA (the radiative heating HRAD is the
  conduction of the mean profile); a zero-mean cosine perturbation in x
  diffuses with unit conductivity (the discrete Laplacian's eigenvalue);
* STEP: the thermal limit uses 1 + RKAPA;
* BCON: the walls take TU, DZTU and DZTB.
Planted mistakes each fail one of these checks.
"""
import math

import numpy as np
import pytest

import oracle

NAD, NRAD0 = float(np.float32(0.41)), float(np.float32(0.4))
BASE = ["--npx", "16", "--npy", "16", "--npz", "64", "--npey", "2", "--npez", "2", "--lrem", "1",
        "--ngrid", "0", "--zmax", "4", "--ampt", "0"]


def kappa(z, rho):
    return rho + (1.25 - rho) * math.exp(-5.0 * z)


def slopes(z, t, rho, nad=NAD, nrad0=NRAD0, kap=kappa):
    """dT/dz and dln p/dz of the background ODE at (z, T, rho)."""
    nrad = nrad0 / kap(z, rho)
    return (nrad if nrad <= nad else nad), 1.0 / t


def radiative(z, t, rho, nad=NAD, nrad0=NRAD0, kap=kappa):
    return nrad0 / kap(z, rho) <= nad


def ode_residual(bg, zs, h=1e-3, **kw):
    """Largest relative mismatch between centred differences of the solution
    bg(z) -> (T, rho) and the ODE's right-hand side, away from the switch."""
    worst = 0.0
    for z in zs:
        (tm, rm), (t0, r0), (tp, rp) = bg(z - h), bg(z), bg(z + h)
        dT = (tp - tm) / (2 * h)
        dlnp = (math.log(tp * rp) - math.log(tm * rm)) / (2 * h)
        sT, sp = slopes(z, t0, r0, **kw)
        # skip points within h of the radiative/adiabatic switch (a kink)
        if radiative(z - h, tm, rm, **kw) != radiative(z + h, tp, rp, **kw):
            continue
        worst = max(worst, abs(dT - sT) / abs(sT), abs(dlnp - sp) / abs(sp))
    return worst


ZS = [0.05, 0.2, 0.5, 0.8, 1.2, 1.7, 2.3, 2.9, 3.5]


def test_background_starts_at_one_and_solves_its_ode():
    assert oracle.lrem_background(2.0) == (1.0, 1.0)
    assert ode_residual(oracle.lrem_background, ZS) < 2e-5


def rk4_solution(z, nad=NAD, nrad0=NRAD0, kap=kappa, dpdz=lambda p, t: p / t, n=4000):
    """A plain fixed-step RK4 solution of an ODE of the same form, for the
    planted mistakes (and a cross-check of the oracle's adaptive one)."""
    t, p, x = 1.0, 1.0, 2.0
    h = (z - 2.0) / n

    def f(x, t, p):
        nrad = nrad0 / kap(x, p / t)
        return (nrad if nrad <= nad else nad), dpdz(p, t)
    for _ in range(n):
        k1 = f(x, t, p)
        k2 = f(x + h / 2, t + h / 2 * k1[0], p + h / 2 * k1[1])
        k3 = f(x + h / 2, t + h / 2 * k2[0], p + h / 2 * k2[1])
        k4 = f(x + h, t + h * k3[0], p + h * k3[1])
        t += h / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
        p += h / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
        x += h
    return t, p / t


def test_background_agrees_with_an_independent_integration():
    for z in (0.3, 1.0, 3.0):
        t, r = oracle.lrem_background(z)
        t2, r2 = rk4_solution(z)
        assert abs(t - t2) < 1e-6 * abs(t2) and abs(r - r2) < 1e-6 * abs(r2), (z, t, t2, r, r2)


@pytest.mark.parametrize("mistake", ["nad_0.4", "nrad_0.41", "dpdz_p_times_t"])
def test_ode_check_catches_planted_mistakes(mistake):
    kw, dpdz = {}, (lambda p, t: p / t)
    if mistake == "nad_0.4":
        kw["nad"] = 0.4
    if mistake == "nrad_0.41":
        kw["nrad0"] = 0.41
    if mistake == "dpdz_p_times_t":
        dpdz = lambda p, t: p * t  # noqa: E731

    def wrong(z):
        return rk4_solution(z, dpdz=dpdz, n=600, **kw)
    # the wrong solution passes its own ODE but not the right one
    assert ode_residual(wrong, ZS) > 1e-3


def test_adiabatic_stretch_is_linear_with_the_polytropic_density():
    # between z = 0.6 and 1.5 the radiative gradient 0.4/kappa exceeds 0.41
    z0, t0, r0 = 0.6, *oracle.lrem_background(0.6)
    assert slopes(z0, t0, r0)[0] == NAD
    for z in (0.8, 1.0, 1.3):
        t, r = oracle.lrem_background(z)
        assert slopes(z, t, r)[0] == NAD
        assert abs(t - (t0 + NAD * (z - z0))) < 1e-7
        p = (t0 * r0) * (t / t0) ** (1.0 / NAD)
        assert abs(t * r - p) < 1e-6 * p


def test_profile_conductivity_and_boundary_values():
    sim = oracle.Sim(BASE)
    d = sim.derived()
    pr = sim.lrem_profile()
    npz = sim.npz
    zmax = 4.0
    z = np.array([float(np.float32(k) / np.float32(npz - 1)) for k in range(npz)]) * zmax
    scale = float(np.float32(8.07)) * 0.25 * d["repr"]
    for k in range(npz):
        t, r = oracle.lrem_background(z[k])
        assert pr["T"][k] == t and pr["R"][k] == r
    check_conductivity(pr["RK"], pr["R"], z, scale)
    # DRK: second-order difference of RK; against the analytic derivative
    # d kappa/dz = rho'(1 - e^(-5z)) - 5 (1.25 - rho) e^(-5z), rho' = rho (1 - T')/T
    dz = zmax / (npz - 1)
    for k in range(2, npz - 2):
        t, r = pr["T"][k], pr["R"][k]
        tp = slopes(z[k], t, r)[0]
        rp = r * (1.0 - tp) / t
        e = math.exp(-5.0 * z[k])
        dk = (rp * (1.0 - e) - 5.0 * (1.25 - r) * e) * scale
        assert abs(pr["DRK"][k] - dk) < 3e-2 * abs(dk) + 2e-3, (k, pr["DRK"][k], dk)
    # TU = T(top); DZTU, DZTB are the one-sided wall gradients of the profile,
    # -/+ (2/3) dz dT/dz to first order: the sign and size of the slope there
    assert pr["TU"] == pr["T"][0]
    s_top = slopes(0.0, pr["T"][0], pr["R"][0])[0]
    s_bot = slopes(zmax, pr["T"][-1], pr["R"][-1])[0]
    assert abs(pr["DZTU"] - (-2.0 / 3.0 * dz * s_top)) < 0.15 * abs(pr["DZTU"])
    assert abs(pr["DZTB"] - (2.0 / 3.0 * dz * s_bot)) < 0.15 * abs(pr["DZTB"])
    # and with them the flux walls hold the profile's own wall values: the
    # one-sided condition T(wall) = c43 T(1) - c13 T(2) + DZT gives T(wall) back
    T = pr["T"]
    assert abs(d["c43"] * T[1] - d["c13"] * T[2] + pr["DZTU"] - T[0]) < 1e-15
    assert abs(d["c43"] * T[-2] - d["c13"] * T[-3] + pr["DZTB"] - T[-1]) < 1e-15
    # the rank metrics carry RK / DRK, ghosts with the end values
    g = np.concatenate([sim.metric(r, "rkapa")[1:-1] for r in (0, 2)])
    assert np.array_equal(g, pr["RK"])
    assert sim.metric(0, "rkapa")[0] == pr["RK"][0] and sim.metric(2, "rkapa")[-1] == pr["RK"][-1]


def check_conductivity(rk, rho, z, scale):
    """RK / (8.07 theta repr) is 1.25 at the top, tends to rho at depth, and
    the excess over rho decays as exp(-5 z)."""
    kap = np.asarray(rk) / scale
    assert abs(kap[0] - 1.25) <= 1e-14
    assert z[0] == 0.0
    for k in range(1, len(z)):
        ex = (kap[k] - rho[k]) / (1.25 - rho[k])
        if z[k] < 2.0:  # below that the excess is under the rounding of rho
            assert abs(math.log(ex) / z[k] + 5.0) < 1e-8, (k, z[k], ex)
    deep = z > 2.5
    assert np.allclose(kap[deep], np.asarray(rho)[deep], rtol=1e-4)


@pytest.mark.parametrize("mistake", ["no_exp", "exp_4z", "top_1", "no_scale"])
def test_conductivity_check_catches_planted_mistakes(mistake):
    sim = oracle.Sim(BASE)
    d = sim.derived()
    pr = sim.lrem_profile()
    z = np.array([float(np.float32(k) / np.float32(sim.npz - 1)) for k in range(sim.npz)]) * 4.0
    scale = float(np.float32(8.07)) * 0.25 * d["repr"]
    r = pr["R"]
    e = np.exp({"exp_4z": -4.0}.get(mistake, -5.0) * z)
    top = 1.0 if mistake == "top_1" else 1.25
    kap = r + (top - r) * e if mistake != "no_exp" else r.copy()
    wrong = kap * (scale if mistake != "no_scale" else 1.0)
    with pytest.raises(AssertionError):
        check_conductivity(wrong, r, z, scale)


STD = [a for a in BASE]
del STD[STD.index("--lrem"):STD.index("--lrem") + 2]


def flux_pair():
    """An LREM run and an ordinary one holding the same state and the same
    conductivity profile, fluid at rest."""
    lr, st = oracle.Sim(BASE), oracle.Sim(STD)
    for r in range(lr.nranks):
        for n in ("ru", "rv", "rw", "ro", "tt"):
            st.field(r, n)[...] = lr.field(r, n)
        for m in ("rkapa", "dkapa"):
            st.metric(r, m)[...] = lr.metric(r, m)
    return lr, st


def fluxes_of_one_stage(sims):
    """FT of FLUXES at the start of a step (one stage), and 1/rho there."""
    ri = [[1.0 / sim.field(r, "ro")[1:-1, 1:-1, 1:-1] for r in range(sim.nranks)] for sim in sims]
    for sim in sims:
        sim.step(1, nstages=1)
    ft = [[sim.field(r, "ft")[1:-1, 1:-1, 1:-1].copy() for r in range(sim.nranks)] for sim in sims]
    return ft, ri


def test_uniform_planes_conduct_like_the_ordinary_diffusion():
    lr, st = flux_pair()
    (fa, fb), _ = fluxes_of_one_stage((lr, st))
    differs = False
    for a, b in zip(fa, fb):
        scale = np.abs(b).max()
        assert scale > 0
        assert np.abs(a - b).max() <= 1e-10 * scale, np.abs(a - b).max() / scale
        differs |= bool((a != b).any())
    assert differs  # a different operation order, the same value


def test_perturbation_diffuses_with_unit_conductivity():
    lr, st = flux_pair()
    amp = 1e-3
    for sim in (lr, st):
        for r in range(sim.nranks):
            i = np.arange(sim.nx)
            # zero mean over the periodic x interior
            sim.field(r, "tt")[...] += amp * np.cos(2 * np.pi * (i - 1) / sim.npx)[None, None, :]
    (fa, fb), (ri, _) = fluxes_of_one_stage((lr, st))
    d = lr.derived()
    lam = -(4.0 * (lr.npx - 1) ** 2 / 20.0 ** 2) * math.sin(math.pi / lr.npx) ** 2
    for r in range(lr.nranks):
        i = np.arange(1, lr.nx - 1)
        cosx = np.cos(2 * np.pi * (i - 1) / lr.npx)[None, None, :]
        rk = lr.metric(r, "rkapa")[1:-1][:, None, None]
        want = ri[r] * (d["ocv"] / d["repr"]) * lam * amp * cosx * (1.0 - rk)
        scale = np.abs(want).max()
        assert scale > 0
        assert np.abs((fa[r] - fb[r]) - want).max() <= 2e-6 * scale, r


def test_thermal_time_step_limit_uses_one_plus_rkapa():
    sim = oracle.Sim(BASE)
    d = sim.derived()
    sim.step_dt()
    want = np.inf
    for r in range(sim.nranks):
        ro = sim.field(r, "ro")[1:-1, 1:-1, 1:-1]
        rk = sim.metric(r, "rkapa")[1:-1][:, None, None]
        dd = min(sim.metric(r, "ddx")[1], sim.metric(r, "ddy")[1], sim.metric(r, "ddz")[1])
        w2 = 0.5 * dd * dd * d["repr"] * d["cv"] * ro / (1.0 + rk)
        assert np.allclose(sim.field(r, "ww2")[1:-1, 1:-1, 1:-1], w2, rtol=1e-15, atol=0)
        want = min(want, float(w2.min()))
    # with RKAPA alone (the mistake) the limit would differ
    ro0 = sim.field(0, "ro")[1:-1, 1:-1, 1:-1]
    rk0 = sim.metric(0, "rkapa")[1:-1][:, None, None]
    assert not np.allclose(ro0 / (1.0 + rk0), ro0 / rk0, rtol=1e-3)


@pytest.mark.parametrize("itcon,izcon", [(0, 0), (1, 1)])
def test_walls_take_the_stratification_boundary_values(itcon, izcon):
    args = BASE + ["--tp", "0", "--itcon", str(itcon), "--izcon", str(izcon)]
    sim = oracle.Sim(args)
    pr = sim.lrem_profile()
    d = sim.derived()
    sim.step(1)
    top = [r for r in range(sim.nranks) if sim.rank_coords(r)[1] == 0]
    bot = [r for r in range(sim.nranks) if sim.rank_coords(r)[1] == sim.npez - 1]
    for r in top:
        tt = sim.field(r, "tt")
        if itcon == 0:
            assert (tt[1, 1:-1, 1:-1] == pr["TU"]).all()
            assert pr["TU"] != 1.0  # not the ordinary constant top temperature
        else:
            want = d["c43"] * tt[2] - d["c13"] * tt[3] + pr["DZTU"]
            assert np.array_equal(tt[1, 1:-1, 1:-1], want[1:-1, 1:-1])
    for r in bot:
        tt = sim.field(r, "tt")
        k = sim.nz - 2
        if izcon == 0:
            assert (tt[k, 1:-1, 1:-1] == pr["T"][-1]).all()
        else:
            want = d["c43"] * tt[k - 1] - d["c13"] * tt[k - 2] + pr["DZTB"]
            assert np.array_equal(tt[k, 1:-1, 1:-1], want[1:-1, 1:-1])


def test_refused_combinations():
    with pytest.raises(oracle.OracleError, match="NGRID"):
        oracle.Sim([a if a != "0" else "1" for a in BASE[:12]] + ["--zmax", "4"]).step(1)
    with pytest.raises(oracle.OracleError, match="LREM"):
        oracle.Sim(BASE + ["--rlax", "0.5", "--itcon", "0", "--izcon", "1", "--toff", "1"]).step(1)
