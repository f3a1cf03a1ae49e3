"""The per-cell flux and its hand-written adjoint (rast-3dmhd_amd/csrc/r3d_adj.hpp),
run on the CPU through tests/adjcheck/adj_check.hip:

* cell_flux over a rank array reproduces the oracle's FLUXES (pinned to the
  reference CPU build) bit for bit, walls and rotation included;
* cell_flux_vjp is the transpose of the flux Jacobian: for random directions
  the adjoint's inner product matches a central finite difference of
  lam . F, field by field, and likewise for the coefficients ore and tk;
* the same for the MHD pair cell_flux_mag / cell_flux_mag_vjp against the
  oracle's LMAG branch, with a fully three-dimensional B;
* cell_flux_mag_extra is the LMAG part of cell_flux_mag (the GPU transpose
  of an MHD stage runs it as a kernel of its own), and every stencil entry
  outside vjp_mask, which the tiled GPU transpose does not store, is zero.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT

ARGS = ["--npx", "12", "--npy", "10", "--npz", "16", "--npey", "1", "--npez", "2",
        "--lrot", "1", "--ry", "0.3", "--ang", "0.4", "--re", "1.7", "--theta", "0.21"]
GRAV = 0.625
X_NAMES = ["ru", "rv", "rw", "ro", "tt"]
F_NAMES = ["fu", "fv", "fw", "fr", "ft"]
MET = ["dzzdz", "d2zzdz2", "rkapa", "dkapa"]


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("adj") / "libadjcheck.so"
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-fPIC",
                    "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "rast-3dmhd_amd", "csrc"), "-shared", "-o", str(out),
                    os.path.join(ROOT, "tests", "adjcheck", "adj_check.hip")],
                   check=True, capture_output=True)
    L = ctypes.CDLL(str(out))
    pp = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))
    dp = ctypes.POINTER(ctypes.c_double)
    L.adj_flux.argtypes = [ctypes.POINTER(ctypes.c_int), pp, pp, dp, pp]
    L.adj_vjp.argtypes = [ctypes.POINTER(ctypes.c_int), pp, pp, dp, pp, pp, dp]
    L.adj_flux_mag.argtypes = [ctypes.POINTER(ctypes.c_int), pp, pp, pp, dp, pp]
    L.adj_vjp_mag.argtypes = [ctypes.POINTER(ctypes.c_int), pp, pp, pp, dp, pp, pp, dp]
    L.adj_flux_mag_extra.argtypes = [ctypes.POINTER(ctypes.c_int), pp, pp, pp, dp, pp]
    L.adj_mask_violations.argtypes = [ctypes.POINTER(ctypes.c_int), pp, pp, pp, dp, pp,
                                      ctypes.c_int]
    L.adj_mask_violations.restype = ctypes.c_long
    L.adj_mag_points_diff.argtypes = [ctypes.POINTER(ctypes.c_int), pp, pp, pp, dp, pp, dp]
    L.adj_mag_points_diff.restype = ctypes.c_long
    return L


def _pp(arrs):
    return (ctypes.POINTER(ctypes.c_double) * len(arrs))(
        *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in arrs])


class Rank:
    def __init__(self, sim, r):
        self.sim, self.r = sim, r
        py, pz = sim.rank_coords(r)
        self.dims = (ctypes.c_int * 6)(sim.nx, sim.ny, sim.nz, int(pz == 0),
                                       int(pz == sim.npez - 1), int(sim.npez > 1))
        self.met = [np.ascontiguousarray(sim.metric(r, n)) for n in
                    ("dxxdx", "d2xxdx2", "dyydy", "d2yydy2", "dzzdz", "d2zzdz2", "rkapa", "dkapa")]
        d = sim.derived()
        self.d = d
        self.k = self.consts(d["ore"], d["ocv"] / d["repr"])

    def consts(self, ore, tk):
        d = self.d
        return np.array([d["c13"], d["c43"], ore, d["ocv"], GRAV, d["hx"], d["h2x"], d["hy"],
                         d["h2y"], d["hz"], d["h2z"], d["omx"], d["omz"], tk, d["ocv"] * ore,
                         d["c13"] * ore, d["c43"] * d["ocv"] * ore, 1.0, d["obeta"], d["orm"],
                         d["obeta"] * d["orm"] * d["ocv"]])

    def flux(self, L, X, k=None):
        F = [np.zeros_like(X[0]) for _ in range(5)]
        kk = self.k if k is None else k
        L.adj_flux(self.dims, _pp(X), _pp(self.met), kk.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)), _pp(F))
        return F

    def flux_mag(self, L, X, B):
        F = [np.zeros_like(X[0]) for _ in range(8)]
        L.adj_flux_mag(self.dims, _pp(X), _pp(B), _pp(self.met),
                       self.k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _pp(F))
        return F

    def vjp_mag(self, L, X, B, lam):
        out = [np.zeros_like(X[0]) for _ in range(8)]
        g = np.zeros(2)
        L.adj_vjp_mag(self.dims, _pp(X), _pp(B), _pp(self.met),
                      self.k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _pp(lam),
                      _pp(out), g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def vjp(self, L, X, lam):
        out = [np.zeros_like(X[0]) for _ in range(5)]
        g = np.zeros(2)
        L.adj_vjp(self.dims, _pp(X), _pp(self.met),
                  self.k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _pp(lam), _pp(out),
                  g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out, g


@pytest.fixture(scope="module")
def sim():
    s = oracle.Sim(ARGS)
    s.step(3)
    return s


def test_cell_flux_reproduces_oracle_fluxes(lib, sim):
    for r in range(sim.nranks):
        rk = Rank(sim, r)
        X = [sim.field(r, n).copy() for n in X_NAMES]
        F = rk.flux(lib, X)
        sim.fluxes(r)
        for f, name in zip(F, F_NAMES):
            want = sim.field(r, name)[1:-1, 1:-1, 1:-1]
            got = f[1:-1, 1:-1, 1:-1]
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (r, name)
        for q, n in enumerate(X_NAMES):  # undo the oracle's in-place 1/RO
            sim.field(r, n)[...] = X[q]


def _interior_mask(shape):
    m = np.zeros(shape)
    m[1:-1, 1:-1, 1:-1] = 1.0
    return m


@pytest.mark.parametrize("field", range(5), ids=X_NAMES)
def test_vjp_matches_finite_differences(lib, sim, field):
    rng = np.random.default_rng(17 + field)
    for r in range(sim.nranks):
        rk = Rank(sim, r)
        X = [sim.field(r, n).copy() for n in X_NAMES]
        mask = _interior_mask(X[0].shape)
        lam = [rng.uniform(-1, 1, X[0].shape) * mask for _ in range(5)]
        scale = np.abs(X[field]).max()
        dirn = rng.uniform(-1, 1, X[0].shape) * scale
        lamX, _ = rk.vjp(lib, X, lam)
        adj = float(np.sum(lamX[field] * dirn))
        # lam . F is quadratic in the momenta (a large step has no truncation
        # error and keeps rounding small); 1/RO makes it curved in RO
        eps = {0: 1e-4, 1: 1e-4, 2: 1e-4, 3: 1e-6, 4: 1e-5}[field]
        Xp = [x.copy() for x in X]
        Xm = [x.copy() for x in X]
        Xp[field] += eps * dirn
        Xm[field] -= eps * dirn
        phi = lambda F: sum(float(np.sum(l * f)) for l, f in zip(lam, F))
        fd = (phi(rk.flux(lib, Xp)) - phi(rk.flux(lib, Xm))) / (2 * eps)
        assert abs(adj - fd) <= 1e-7 * max(abs(fd), 1e-300) + 1e-12, (r, X_NAMES[field], adj, fd)


def test_vjp_coefficient_derivatives(lib, sim):
    rng = np.random.default_rng(5)
    for r in range(sim.nranks):
        rk = Rank(sim, r)
        X = [sim.field(r, n).copy() for n in X_NAMES]
        mask = _interior_mask(X[0].shape)
        lam = [rng.uniform(-1, 1, X[0].shape) * mask for _ in range(5)]
        _, g = rk.vjp(lib, X, lam)
        phi = lambda F: sum(float(np.sum(l * f)) for l, f in zip(lam, F))
        ore, tk = rk.d["ore"], rk.d["ocv"] / rk.d["repr"]
        # lam . F is linear in ore and tk: a large step is exact up to rounding
        for idx, (vo, vt) in enumerate([((1 + 1e-3) * ore, tk), (ore, (1 + 1e-3) * tk)]):
            lo = (2 * ore - vo, 2 * tk - vt)
            fp = phi(rk.flux(lib, X, rk.consts(vo, vt)))
            fm = phi(rk.flux(lib, X, rk.consts(*lo)))
            h = (vo - ore) if idx == 0 else (vt - tk)
            fd = (fp - fm) / (2 * h)
            assert abs(g[idx] - fd) <= 1e-7 * abs(fd) + 1e-12, (r, idx, g[idx], fd)


MAG_ARGS = ["--npx", "12", "--npy", "10", "--npz", "18", "--npey", "1", "--npez", "2",
            "--lmag", "1", "--rm", "5", "--beta", "2", "--ampb", "1", "--bfh", "2",
            "--bzp", "10", "--lrot", "1", "--ry", "0.3", "--ang", "0.4"]
B_NAMES = ["bx", "by", "bz"]


@pytest.fixture(scope="module")
def msim():
    """An MHD state with a fully three-dimensional B (the layer alone has
    by = bz = 0)."""
    s = oracle.Sim(MAG_ARGS)
    s.step(2)
    rng = np.random.default_rng(9)
    for r in range(s.nranks):
        for n in B_NAMES:
            f = s.field(r, n)
            f[1:-1, 1:-1, 1:-1] += 0.2 * rng.uniform(-1, 1, f[1:-1, 1:-1, 1:-1].shape)
    s.communicate()
    return s


def test_cell_flux_mag_reproduces_oracle_fluxes(lib, msim):
    for r in range(msim.nranks):
        rk = Rank(msim, r)
        X = [msim.field(r, n).copy() for n in X_NAMES]
        B = [msim.field(r, n).copy() for n in B_NAMES]
        F = rk.flux_mag(lib, X, B)
        msim.fluxes(r)
        for f, name in zip(F, F_NAMES + ["ww1", "ww2", "ww3"]):
            want = msim.field(r, name)[1:-1, 1:-1, 1:-1]
            got = f[1:-1, 1:-1, 1:-1]
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (r, name)
        for q, n in enumerate(X_NAMES):
            msim.field(r, n)[...] = X[q]


@pytest.mark.parametrize("field", range(8), ids=X_NAMES + B_NAMES)
def test_vjp_mag_matches_finite_differences(lib, msim, field):
    rng = np.random.default_rng(31 + field)
    names = X_NAMES + B_NAMES
    for r in range(msim.nranks):
        rk = Rank(msim, r)
        XB = [msim.field(r, n).copy() for n in names]
        mask = _interior_mask(XB[0].shape)
        lam = [rng.uniform(-1, 1, XB[0].shape) * mask for _ in range(8)]
        dirn = rng.uniform(-1, 1, XB[0].shape) * np.abs(XB[field]).max()
        lamX = rk.vjp_mag(lib, XB[:5], XB[5:], lam)
        adj = float(np.sum(lamX[field] * dirn))
        # lam . F is quadratic in the momenta and in B; 1/RO curves it in RO
        eps = {3: 1e-6, 4: 1e-5}.get(field, 1e-4)
        P = [x.copy() for x in XB]
        M = [x.copy() for x in XB]
        P[field] += eps * dirn
        M[field] -= eps * dirn
        phi = lambda F: sum(float(np.sum(l * f)) for l, f in zip(lam, F))
        fd = (phi(rk.flux_mag(lib, P[:5], P[5:])) - phi(rk.flux_mag(lib, M[:5], M[5:]))) / (2 * eps)
        assert abs(adj - fd) <= 1e-7 * abs(fd) + 1e-12, (r, names[field], adj, fd)


def test_lmag_terms_are_the_difference_of_the_fluxes(lib, msim):
    for r in range(msim.nranks):
        rk = Rank(msim, r)
        X = [msim.field(r, n).copy() for n in X_NAMES]
        B = [msim.field(r, n).copy() for n in B_NAMES]
        Fm = rk.flux_mag(lib, X, B)
        Fh = rk.flux(lib, X)
        E = [np.zeros_like(X[0]) for _ in range(8)]
        lib.adj_flux_mag_extra(rk.dims, _pp(X), _pp(B), _pp(rk.met),
                               rk.k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _pp(E))
        sl = (slice(1, -1),) * 3
        for q in range(8):
            want = Fm[q][sl] - (Fh[q][sl] if q < 5 else 0.0)
            # rounding of the hydrodynamic part, relative to the whole flux
            tol = 1e-13 * (np.abs(Fm[q][sl]).max() + 1e-300)
            assert np.abs(E[q][sl] - want).max() <= tol, (r, q)


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["hydro", "mhd", "lmag_terms"])
def test_transpose_is_zero_outside_the_stored_mask(lib, sim, msim, mode):
    s = msim if mode else sim
    rng = np.random.default_rng(77 + mode)
    names = X_NAMES + (B_NAMES if mode else [])
    for r in range(s.nranks):
        rk = Rank(s, r)
        X = [s.field(r, n).copy() for n in X_NAMES]
        B = [s.field(r, n).copy() for n in B_NAMES] if mode else X[:3]
        mask = _interior_mask(X[0].shape)
        lam = [rng.uniform(-1, 1, X[0].shape) * mask for _ in range(8)]
        bad = lib.adj_mask_violations(rk.dims, _pp(X), _pp(B), _pp(rk.met),
                                      rk.k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      _pp(lam), mode)
        assert bad == 0, (r, names, bad)


def test_point_major_lmag_transpose_equals_term_major(lib, msim):
    """cell_flux_mag_extra_vjp_points (the GPU transpose's form: each face
    point's entries in closed form, emitted once) against the term-major
    cell_flux_mag_extra_vjp, which the finite-difference checks above pin:
    the same entries up to the order of a few sums, on every rank, walls with
    their one-sided stencils included."""
    rng = np.random.default_rng(123)
    walls = 0.0
    for r in range(msim.nranks):
        rk = Rank(msim, r)
        X = [msim.field(r, n).copy() for n in X_NAMES]
        B = [msim.field(r, n).copy() for n in B_NAMES]
        mask = _interior_mask(X[0].shape)
        lam = [rng.uniform(-1, 1, X[0].shape) * mask for _ in range(8)]
        out = np.zeros(4)
        bad = lib.adj_mag_points_diff(rk.dims, _pp(X), _pp(B), _pp(rk.met),
                                      rk.k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      _pp(lam), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert bad == 0, r
        assert out[0] <= 1e-13 * out[2], (r, out)
        assert out[1] <= 1e-13 * max(out[3], 1e-300), (r, out)
        walls = max(walls, out[3])
    assert walls > 0.0  # some rank holds a wall with one-sided LMAG stencils
