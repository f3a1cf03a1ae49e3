"""Flux relaxation RLAX (M. Rempel), ITC=0 / IZC=1 (CPU: oracle and host).

golden/fortran/3dmhdsub.f:1534-1547, 1625-1771: per FLUXES call the
horizontal means TTM of T and FCONM of WW2 = -RW*(2.5*(T-TTM)+0.5*|u|**2)
give ALPHA = a/(1+4|a|), a = 1-(FCONM+FRADM)/FTOT with FRADM = dTTM/dz *
RKAPA/REPR and FTOT = THETA/REPR; CORRECT is the trapezoid integral of ALPHA
from the top wall, continued across the z ranks, and FT gains
RLAX*TFAC*CORRECT(K).  TFAC rises by DT/(3 TOFF) (at most to 2) while the
bottom-wall TTM sets a new maximum and falls by as much otherwise.  The
reference C++ refuses RLAX (src/fluxes.cpp:48): parity unpinned, pinned here
by closed forms on uniform grids written differently from the oracle:

* T linear in z, RW = 0: ALPHA is the constant 1 - s/THETA (mapped), so
  CORRECT = ALPHA*(z - z_top) across all z ranks (the rank offsets included);
* RW = e cos(th i), T = T(z) + d cos(th i): FCONM = -1.25 e d, the 2.5*(T-TTM)
  term alone (the cubic kinetic term averages out over whole periods);
* RW uniform: FCONM = -0.5 RW**3/RO**2;
* TFAC = 1 + DT/(3 TOFF) after the first call, then one step down when the
  bottom-wall mean does not rise;
* FT changes by exactly RLAX*TFAC*CORRECT (the difference to RLAX = 0).
"""
import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()

GRID = ["--npx", "12", "--npy", "10", "--npz", "18", "--npey", "2", "--npez", "3",
        "--ngrid", "0"]
RLX = ["--itcon", "0", "--izcon", "1", "--rlax", "0.7", "--tstart", "-1", "--toff", "40"]
RHO = 1.3


def fill(sim, tt_of, rw_of):
    nry, nrz = sim.npy // sim.npey, sim.npz // sim.npez
    for r in range(sim.nranks):
        py, pz = sim.rank_coords(r)
        k, j, i = np.meshgrid(pz * nrz + np.arange(sim.nz), py * nry + np.arange(sim.ny),
                              np.arange(sim.nx), indexing="ij")
        z = sim.metric(r, "zee")[:, None, None] * np.ones_like(i, dtype=float)
        sim.field(r, "ru")[...] = 0.0
        sim.field(r, "rv")[...] = 0.0
        sim.field(r, "ro")[...] = RHO
        sim.field(r, "tt")[...] = tt_of(i, z)
        sim.field(r, "rw")[...] = rw_of(i, z)


def expected_correct(sim, alpha):
    """alpha mapped by a/(1+4|a|), integrated from the top wall."""
    a = alpha / (1.0 + 4.0 * abs(alpha))
    z_top = sim.metric(0, "zee")[1]
    return [a * (sim.metric(r, "zee")[1:-1] - z_top) for r in range(sim.nranks)]


def check(sim, want, rtol=1e-7):
    # the uniform grid's derivative factors carry the single-precision spacing
    # f4(1/(npz-1)) of the reference grid, a relative 1e-8 on FRADM
    scale = max(np.abs(w).max() for w in want)
    assert scale > 1e-3
    for r in range(sim.nranks):
        np.testing.assert_allclose(sim.correct(r)[1:-1], want[r], rtol=0, atol=rtol * scale)


def test_linear_temperature_gives_a_linear_correction_across_ranks():
    sim = oracle.Sim(GRID + RLX)
    s, theta = 0.15, 0.25  # THETA: the default
    fill(sim, lambda i, z: 1.0 + s * z, lambda i, z: 0.0 * z)
    dt = 0.37  # the step's dt as FLUXES sees it (no step has been taken)
    sim.set_scalars(dict(sim.scalars(), dt=dt))
    sim.relax_prep()
    check(sim, expected_correct(sim, 1.0 - s / theta))
    st = sim.relax_state()
    assert st["tfac"] == pytest.approx(1.0 + dt / (3.0 * 40.0), rel=1e-15)
    assert st["ttm_max"] == pytest.approx(1.0 + s * sim.metric(sim.nranks - 1, "zee")[-2], rel=1e-14)
    # the same state again: no new maximum, TFAC steps down
    t1 = st["tfac"]
    sim.relax_prep()
    assert sim.relax_state()["tfac"] == pytest.approx(t1 - dt / (3.0 * 40.0), rel=1e-15)
    # a rising bottom-wall mean raises TFAC again, capped at 2
    fill(sim, lambda i, z: 1.5 + s * z, lambda i, z: 0.0 * z)
    sim.set_scalars(dict(sim.scalars(), dt=400.0))
    sim.relax_prep()
    assert sim.relax_state()["tfac"] == 2.0


def test_convective_flux_terms():
    th = 2 * np.pi * 2 / 12
    e, d, s, theta = 0.02, 0.3, 0.1, 0.25
    repr_ = oracle.Sim(GRID + RLX).derived()["repr"]
    # the 2.5*(T - TTM) term: FCONM = -1.25 e d
    sim = oracle.Sim(GRID + RLX)
    fill(sim, lambda i, z: 1.0 + s * z + d * np.cos(th * i), lambda i, z: e * np.cos(th * i) + 0 * z)
    sim.relax_prep()
    check(sim, expected_correct(sim, 1.0 - (-1.25 * e * d + s / repr_) / (theta / repr_)))
    # uniform RW: FCONM = -0.5 RW**3 / RO**2
    c = 0.05
    sim = oracle.Sim(GRID + RLX)
    fill(sim, lambda i, z: 1.0 + s * z, lambda i, z: c + 0 * z)
    sim.relax_prep()
    check(sim, expected_correct(sim, 1.0 - (-0.5 * c ** 3 / RHO ** 2 + s / repr_) / (theta / repr_)))


def test_ft_gains_rlax_tfac_correct():
    s = 0.15
    out = {}
    for tag, extra in (("on", RLX), ("off", ["--itcon", "0", "--izcon", "1"])):
        sim = oracle.Sim(GRID + extra)
        fill(sim, lambda i, z: 1.0 + s * z, lambda i, z: 0.0 * z)
        if tag == "on":
            sim.relax_prep()
            tfac = sim.relax_state()["tfac"]
            cor = [sim.correct(r).copy() for r in range(sim.nranks)]
        out[tag] = []
        for r in range(sim.nranks):
            sim.fluxes(r)
            out[tag].append(sim.field(r, "ft")[1:-1, 1:-1, 1:-1].copy())
    for r in range(len(cor)):
        want = 0.7 * tfac * cor[r][1:-1][:, None, None]
        diff = out["on"][r] - out["off"][r]
        np.testing.assert_allclose(diff, np.broadcast_to(want, diff.shape), rtol=0,
                                   atol=1e-12 * max(1.0, np.abs(out["off"][r]).max()))


def test_relaxation_flags_are_checked():
    assert pkg.Run(GRID + RLX).cfg().rlax == 0.7
    with pytest.raises(pkg.HipError, match="RLAX"):
        pkg.Run(GRID + ["--rlax", "0.7", "--toff", "40"])  # ITCON=1 / IZCON=0
    with pytest.raises(pkg.HipError, match="NGRID"):
        pkg.Run(GRID[:-2] + RLX)
    with pytest.raises(RuntimeError, match="RLAX"):
        oracle.Sim(GRID + ["--rlax", "0.7", "--toff", "40"]).step(1)
