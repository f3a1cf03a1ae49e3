"""The oracle's LMAG branch (Fortran FLUXES, golden/fortran/3dmhdsub.f:2146-2651).

No reference build runs LMAG (the C++ aborts, src/fluxes.cpp:539-541), so the
restatement is pinned by what the mathematics fixes instead:

* B = 0: the magnetic branch must leave FU..FT bit-identical to the
  hydrodynamic fluxes and produce WW1..WW3 = 0 exactly.
* Polynomial fields on a uniform grid: centred second-order differences are
  exact for quadratics, so the LMAG contributions must equal the continuum
  terms evaluated from the polynomial coefficients: Lorentz force
  obeta (curl B) x B, Ohmic heating obeta orm ocv |curl B|^2 / rho, and the
  induction flux curl(u x B) - u div(B) + orm lap(B) (the Fortran drops the
  u div B term of curl(u x B)).  The vertical velocity is made to vanish on the
  two walls, where the Fortran's one-sided stencil for d(WW)/dz is exact then
  (its one-sided second derivatives of BX, BY are exact for cubics).
  A wrong sign, index, component or factor in any term fails this.
"""
import numpy as np
import pytest

import oracle

BASE = ["--ngrid", "0", "--npx", "12", "--npy", "12", "--npz", "24", "--npey", "2",
        "--npez", "2"]
MAG = ["--lmag", "1", "--rm", "5", "--beta", "2", "--bfh", "2", "--bzp", "10"]


def test_zero_field_reduces_to_hydro():
    hyd = oracle.Sim(BASE)
    mag = oracle.Sim(BASE + MAG + ["--ampb", "0"])
    for r in range(hyd.nranks):
        for name in ("ru", "rv", "rw", "ro", "tt"):
            assert np.array_equal(hyd.field(r, name), mag.field(r, name))
        assert not mag.field(r, "bx").any()
        hyd.fluxes(r)
        mag.fluxes(r)
        for name in ("fu", "fv", "fw", "fr", "ft"):
            a, b = hyd.field(r, name), mag.field(r, name)
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (r, name)
        for name in ("ww1", "ww2", "ww3"):
            w = mag.field(r, name)[1:-1, 1:-1, 1:-1]
            assert not w.any(), (r, name)


class Quad:
    """f(x) = c + g.x + x.A.x (A symmetric) with its exact derivatives."""

    def __init__(self, rng, scale=1.0):
        self.c = rng.uniform(-1, 1) * scale
        self.g = rng.uniform(-1, 1, 3) * scale * 0.3
        a = rng.uniform(-1, 1, (3, 3)) * scale * 0.05
        self.A = 0.5 * (a + a.T)

    def __call__(self, X):
        return self.c + X @ self.g + np.einsum("...i,ij,...j->...", X, self.A, X)

    def grad(self, X):
        return self.g + 2.0 * X @ self.A

    def lap(self):
        return 2.0 * np.trace(self.A)


class Wall:
    """a (z - z1)(z - z2): the only quadratic that vanishes on both walls."""

    def __init__(self, rng, z1, z2):
        self.a, self.z1, self.z2 = rng.uniform(0.5, 1.0) / (z2 - z1) ** 2, z1, z2

    def __call__(self, X):
        z = X[..., 2]
        return self.a * (z - self.z1) * (z - self.z2)

    def grad(self, X):
        g = np.zeros(X.shape)
        g[..., 2] = self.a * (2.0 * X[..., 2] - self.z1 - self.z2)
        return g


def _coords(sim, r):
    """Uniformly extended coordinates (ghosts included) of rank r, shape (nz,ny,nx,3)."""
    d = sim.derived()
    dx = 1.0 / (2.0 * d["hx"] * sim.metric(r, "dxxdx")[1])
    dy = 1.0 / (2.0 * d["hy"] * sim.metric(r, "dyydy")[1])
    dz = 1.0 / (2.0 * d["hz"] * sim.metric(r, "dzzdz")[1])
    py, pz = sim.rank_coords(r)
    nry, nrz = sim.npy // sim.npey, sim.npz // sim.npez
    x = np.arange(sim.nx) * dx
    y = (py * nry + np.arange(sim.ny)) * dy
    z = (pz * nrz + np.arange(sim.nz)) * dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.stack([X, Y, Z], axis=-1), dz


def _curl(J):
    """curl from a Jacobian J[..., comp, dir]."""
    return np.stack([J[..., 2, 1] - J[..., 1, 2], J[..., 0, 2] - J[..., 2, 0],
                     J[..., 1, 0] - J[..., 0, 1]], axis=-1)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_polynomial_fields_match_continuum_terms(seed):
    rng = np.random.default_rng(seed)
    hyd = oracle.Sim(BASE)
    mag = oracle.Sim(BASE + MAG)
    d = mag.derived()
    obeta, orm, ocv = d["obeta"], d["orm"], d["ocv"]
    _, dz = _coords(mag, 0)
    z1, z2 = 1 * dz, mag.npz * dz  # the two wall planes
    vel = [Quad(rng), Quad(rng), Wall(rng, z1, z2)]
    bq = [Quad(rng, 2.0) for _ in range(3)]
    tq = Quad(rng, 0.2)
    worst = 0.0
    for r in range(mag.nranks):
        X, _ = _coords(mag, r)
        for sim in (hyd, mag):
            for c, name in enumerate(("ru", "rv", "rw")):
                sim.field(r, name)[...] = vel[c](X)
            sim.field(r, "ro")[...] = 1.0
            sim.field(r, "tt")[...] = 2.0 + tq(X)
        for c, name in enumerate(("bx", "by", "bz")):
            mag.field(r, name)[...] = bq[c](X)
        hyd.fluxes(r)
        mag.fluxes(r)

        Xi = X[1:-1, 1:-1, 1:-1]
        u = np.stack([v(Xi) for v in vel], axis=-1)
        Ju = np.stack([v.grad(Xi) for v in vel], axis=-2)
        B = np.stack([b(Xi) for b in bq], axis=-1)
        JB = np.stack([b.grad(Xi) for b in bq], axis=-2)
        lapB = np.array([b.lap() for b in bq])
        curlB = _curl(JB)
        divB = np.trace(JB, axis1=-2, axis2=-1)
        divu = np.trace(Ju, axis1=-2, axis2=-1)
        # curl(u x B) = u div B - B div u + (B.grad) u - (u.grad) B
        curl_uxB = (u * divB[..., None] - B * divu[..., None] + np.einsum("...ij,...j->...i", Ju, B)
                    - np.einsum("...ij,...j->...i", JB, u))
        lorentz = obeta * np.cross(curlB, B)
        ohmic = obeta * orm * ocv * np.sum(curlB * curlB, axis=-1)  # rho = 1
        induction = curl_uxB - u * divB[..., None] + orm * lapB

        def inner(sim, name):
            return sim.field(r, name)[1:-1, 1:-1, 1:-1]

        got = {
            "fu": inner(mag, "fu") - inner(hyd, "fu"),
            "fv": inner(mag, "fv") - inner(hyd, "fv"),
            "fw": inner(mag, "fw") - inner(hyd, "fw"),
            "ft": inner(mag, "ft") - inner(hyd, "ft"),
            "ww1": inner(mag, "ww1"), "ww2": inner(mag, "ww2"), "ww3": inner(mag, "ww3"),
        }
        want = {"fu": lorentz[..., 0], "fv": lorentz[..., 1], "fw": lorentz[..., 2],
                "ft": ohmic, "ww1": induction[..., 0], "ww2": induction[..., 1],
                "ww3": induction[..., 2]}
        for k in got:
            scale = max(np.abs(want[k]).max(), 1e-3)
            err = np.abs(got[k] - want[k]).max() / scale
            worst = max(worst, err)
            assert err < 1e-9, (r, k, err)
    assert worst < 1e-9
