"""Density diffusion ID=1/2 (CPU: oracle and host setup).

The Fortran FLUXES diffuses density perturbations (golden/fortran/3dmhdsub.f:
1306-1380): the outer RO ghost planes take their mirror points, ROM is the
horizontal mean of RO (HORIZONTAL_MEAN), WW1 = RO - ROM, and
  FR += (Lx(WW1) + Ly(WW1)) * DH*ORE*WWZ   and   FR += Lz(WW1) * DV*ORE*WWZ
with the metric second-difference operators L and WWZ = 1/ROM (ID=1) or two
Gaussians of width 0.1 at the top and bottom (ID=2).  The reference C++
refuses ID (src/fluxes.cpp:49), so the flux term is parity unpinned; it is
pinned here by closed forms written differently from the oracle:

* the horizontal mean is the reference's own: its horizontal_mean
  (src/communicate.cpp:184-223), built from the reference sources into
  oracle/_ref/hmean, gives bit-identical means on several decompositions;
* on a fluid at rest with RO = R + A cos(theta i) + B cos(phi j), FR is
  the discrete Laplacian eigen-form DH*ORE/ROM*(lambda_x A cos + lambda_y B cos)
  with lambda from the trigonometric identities (the horizontal mean of a
  full cosine period is zero);
* RO = R + A cos(theta i) g(k) with g quadratic in the plane index gives the
  vertical term from exact second differences, and on the two wall planes
  the mirror (g(0) -> g(2), g(npz+1) -> g(npz-1));
* ID=2 weights the same terms with 2**(-4 (z/0.1)**2) + 2**(-4 ((z-zmax)/0.1)**2).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import load_package

pkg = load_package()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HMEAN = os.path.join(ROOT, "oracle", "_ref", "hmean")
MPIRUN = "/opt/conda/bin/mpirun"


def hash_fill(sim, name):
    """The fill of oracle/ref_hmean.cpp, by global ghost-inclusive position."""
    nry, nrz = sim.npy // sim.npey, sim.npz // sim.npez
    for r in range(sim.nranks):
        py, pz = sim.rank_coords(r)
        k, j, i = np.meshgrid(np.arange(sim.nz), np.arange(sim.ny), np.arange(sim.nx),
                              indexing="ij")
        h = (i * 7919 + (py * nry + j) * 104729 + (pz * nrz + k) * 1299709) % 10007
        sim.field(r, name)[...] = 1.0 + h / 10007.0 / 2.0


@pytest.mark.parametrize("npey,npez", [(1, 2), (2, 2), (4, 2), (3, 1)])
def test_horizontal_mean_is_the_reference_one(npey, npez):
    if not (os.path.exists(HMEAN) and os.path.exists(MPIRUN)):
        pytest.skip("reference build (oracle/_ref/hmean) or mpirun not available")
    npx, npy, npz = 13, 12, 8
    out = subprocess.run([MPIRUN, "-np", str(npey * npez), HMEAN, str(npx), str(npy), str(npz),
                          str(npey)], env=dict(os.environ, OMP_NUM_THREADS="1"),
                         capture_output=True, text=True, check=True, timeout=120).stdout
    want = {}
    for line in out.splitlines():
        if line.startswith("HMEAN "):
            _, pz, k, v = line.split()
            want[(int(pz), int(k))] = float.fromhex(v)
    sim = oracle.Sim(["--npx", str(npx), "--npy", str(npy), "--npz", str(npz), "--npey",
                      str(npey), "--npez", str(npez), "--ngrid", "0"])
    hash_fill(sim, "ro")
    sim.hmean("ro")
    assert len(want) == npez * sim.nz
    for r in range(sim.nranks):
        pz = sim.rank_coords(r)[1]
        got = sim.rom(r)
        for k in range(sim.nz):
            assert got[k] == want[(pz, k)], (r, k, got[k].hex(), want[(pz, k)].hex())


BASE = ["--npx", "12", "--npy", "10", "--npz", "16", "--npey", "2", "--npez", "2",
        "--ngrid", "0"]
RHO = 1.7  # mean density of the closed-form states


def at_rest(sim, ro_of):
    """u = 0, T = 1.3, RO from ro_of(ig, jg, kg) on every rank (ghosts included)."""
    nry, nrz = sim.npy // sim.npey, sim.npz // sim.npez
    for r in range(sim.nranks):
        py, pz = sim.rank_coords(r)
        k, j, i = np.meshgrid(pz * nrz + np.arange(sim.nz), py * nry + np.arange(sim.ny),
                              np.arange(sim.nx), indexing="ij")
        for n in ("ru", "rv", "rw"):
            sim.field(r, n)[...] = 0.0
        sim.field(r, "tt")[...] = 1.3
        sim.field(r, "ro")[...] = ro_of(i, j, k)


def fr_after_fluxes(sim):
    sim.density_prep()
    out = []
    for r in range(sim.nranks):
        sim.fluxes(r)
        out.append(sim.field(r, "fr")[1:-1, 1:-1, 1:-1].copy())
    return out


def wwz2(z, zmax):
    return 2.0 ** (-4.0 * (z / 0.1) ** 2) + 2.0 ** (-4.0 * ((z - zmax) / 0.1) ** 2)


@pytest.mark.parametrize("idv", [1, 2])
def test_horizontal_diffusion_is_the_discrete_laplacian(idv):
    args = BASE + ["--id", str(idv), "--dh", "0.3", "--dv", "0", "--zmax", "0.6"]
    sim = oracle.Sim(args)
    A, B, m, q = 0.02, 0.03, 2, 1
    th, ph = 2 * np.pi * m / sim.npx, 2 * np.pi * q / sim.npy
    at_rest(sim, lambda i, j, k: RHO + A * np.cos(th * i) + B * np.cos(ph * j))
    fr = fr_after_fluxes(sim)
    d = sim.derived()
    nry, nrz = sim.npy // sim.npey, sim.npz // sim.npez
    for r in range(sim.nranks):
        py, pz = sim.rank_coords(r)
        i = np.arange(1, sim.nx - 1)
        jg = py * nry + np.arange(1, sim.ny - 1)
        dxx, d2xx = sim.metric(r, "dxxdx")[1:-1], sim.metric(r, "d2xxdx2")[1:-1]
        dyy, d2yy = sim.metric(r, "dyydy")[1:-1], sim.metric(r, "d2yydy2")[1:-1]
        lx = (d["hx"] * d2xx * (-2 * A * np.sin(th * i) * np.sin(th)) +
              d["h2x"] * dxx ** 2 * (-4 * A * np.sin(th / 2) ** 2 * np.cos(th * i)))
        ly = (d["hy"] * d2yy * (-2 * B * np.sin(ph * jg) * np.sin(ph)) +
              d["h2y"] * dyy ** 2 * (-4 * B * np.sin(ph / 2) ** 2 * np.cos(ph * jg)))
        if idv == 1:
            w = np.full(sim.nz - 2, 1.0 / RHO)  # 1/ROM, ROM = RHO (full cosine periods)
        else:
            w = wwz2(sim.metric(r, "zee")[1:-1], 0.6)
        want = 0.3 * d["ore"] * w[:, None, None] * (lx[None, None, :] + ly[None, :, None])
        scale = np.abs(want).max()
        assert scale > 0
        # ID=2: CLN carries the single-precision LOG(2) (relative ~1e-7 in WWZ)
        np.testing.assert_allclose(fr[r], want, rtol=0 if idv == 1 else 2e-6,
                                   atol=1e-11 * max(scale, 1e-3))


@pytest.mark.parametrize("idv", [1, 2])
def test_vertical_diffusion_and_the_mirrored_walls(idv):
    args = BASE + ["--id", str(idv), "--dh", "0", "--dv", "0.2", "--zmax", "0.6"]
    sim = oracle.Sim(args)
    A, m = 0.05, 1
    th = 2 * np.pi * m / sim.npx
    npz = sim.npz

    def g(k):
        return (k / npz) ** 2

    at_rest(sim, lambda i, j, k: RHO + A * np.cos(th * i) * g(k))
    fr = fr_after_fluxes(sim)
    d = sim.derived()
    nrz = npz // sim.npez
    for r in range(sim.nranks):
        pz = sim.rank_coords(r)[1]
        kg = pz * nrz + np.arange(1, sim.nz - 1)
        kp = np.where(kg == npz, npz - 1, kg + 1)  # mirror below the bottom wall
        km = np.where(kg == 1, 2, kg - 1)          # mirror above the top wall
        dzz, d2zz = sim.metric(r, "dzzdz")[1:-1], sim.metric(r, "d2zzdz2")[1:-1]
        lz = d["hz"] * d2zz * (g(kp) - g(km)) + d["h2z"] * dzz ** 2 * (g(kp) - 2 * g(kg) + g(km))
        i = np.arange(1, sim.nx - 1)
        w = np.full(sim.nz - 2, 1.0 / RHO) if idv == 1 else wwz2(sim.metric(r, "zee")[1:-1], 0.6)
        want = 0.2 * d["ore"] * (w * lz)[:, None, None] * A * np.cos(th * i)[None, None, :]
        scale = np.abs(want).max()
        assert scale > 0
        np.testing.assert_allclose(fr[r], np.broadcast_to(want, fr[r].shape),
                                   rtol=0 if idv == 1 else 2e-6, atol=1e-11 * max(scale, 1e-3))
    # the wall planes really see the mirror: their term is not the one-sided value
    top = [r for r in range(sim.nranks) if sim.rank_coords(r)[1] == 0][0]
    assert np.abs(fr[top][0]).max() > 0


def test_id_flags_are_checked():
    assert pkg.Run(BASE + ["--id", "1", "--dh", "0.1"]).cfg().id == 1
    with pytest.raises(pkg.HipError, match="NGRID"):
        pkg.Run(BASE[:-2] + ["--id", "1", "--dh", "0.1"])
    with pytest.raises(pkg.HipError, match="ID=3"):
        pkg.Run(BASE + ["--id", "3"])
    with pytest.raises(RuntimeError, match="ID=3"):
        oracle.Sim(BASE + ["--id", "3"]).step(1)
