"""The last stage's time-step screen (rast-3dmhd_amd/csrc/r3d_dtlim.hpp), on
the CPU through tests/dtcheck/dt_screen_check.hip.

dt_screen decides, with division-free bounds, whether a cell may set an
extremum of the next reduction; the stage kernel evaluates the exact limits
(dt_terms, reference src/step.cpp:125-209) only there, and k_finish accepts
the screened extrema only if screen_proven holds, else the step redoes them
exactly.  Checked here:

* no cell whose exact limits reach the thresholds is passed over, on random
  states and on adversarial ones: thresholds placed on a cell's own limit to
  the last ulp (each limit in turn), cells where the sound-speed bound is
  tight (UU = G/16), MHD with strong fields, cells at rest;
* unphysical values (rho, T or the conductivity <= 0, NaN) are always
  evaluated, and a cell at rest does not pass the umach test at um_prev = 0;
* over random thresholds on both sides of the extrema, whenever
  screen_proven holds the screened reduction equals the exact one, and the
  rule does refuse (a redo) when the thresholds are too tight;
* the adversarial cells are strong enough to catch plausible mistakes in the
  screen: a numpy screen with the sound bound fast <= UU + G (too small) or
  without the margin misses some of them.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

GAMMA = 5.0 / 3.0
K = np.array([GAMMA, 1.0 / GAMMA, 0.1, 1.5, 2.0, 50.0, 0.1, 1.05, 0.95])


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        hipcc = shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("dts") / "libdtscreen.so"
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-fPIC",
                    "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "rast-3dmhd_amd", "csrc"), "-shared", "-o", str(out),
                    os.path.join(ROOT, "tests", "dtcheck", "dt_screen_check.hip")],
                   check=True, capture_output=True)
    L = ctypes.CDLL(str(out))
    dp = ctypes.POINTER(ctypes.c_double)
    L.screen_cells.argtypes = [dp, ctypes.c_long, dp, ctypes.c_int, ctypes.c_double,
                               ctypes.c_double, dp, ctypes.POINTER(ctypes.c_int)]
    L.screen_cells.restype = ctypes.c_long
    L.screen_proven_c.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double]
    return L


def screen(L, cells, mag, lim_prev, um_prev, k=K):
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    n = len(cells)
    exact = np.zeros((n, 5))
    hit = np.zeros(n, dtype=np.int32)
    missed = L.screen_cells(cells.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                            k.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(mag),
                            lim_prev, um_prev, exact.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            hit.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return missed, exact, hit.astype(bool)


def random_cells(rng, n, mag):
    ro = rng.uniform(0.5, 40.0, n)
    c = np.empty((n, 10))
    c[:, 0:3] = rng.normal(0.0, 0.3, (n, 3)) * ro[:, None]
    c[:, 3] = ro
    c[:, 4] = rng.uniform(0.2, 10.0, n)
    c[:, 5:8] = rng.normal(0.0, 2.0, (n, 3)) if mag else 0.0
    c[:, 8] = rng.uniform(0.005, 0.05, n)
    c[:, 9] = rng.uniform(0.5, 2.0, n)
    return c


def tight_sound_cells(rng, n, mag):
    """|u|^2 = G/16, where 2 sqrt(UU G) <= 4 UU + G/4 holds with equality."""
    c = random_cells(rng, n, mag)
    ro, tt = c[:, 3], c[:, 4]
    g = GAMMA * tt
    if mag:
        g = g + (c[:, 5:8] ** 2).sum(1) * K[6] / ro
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    c[:, 0:3] = d * (np.sqrt(g / 16.0) * ro)[:, None]
    return c


def limits(exact, mag):
    return exact[:, 1:5] if mag else exact[:, 1:4]


@pytest.mark.parametrize("mag", [False, True], ids=["hydro", "mhd"])
def test_no_cell_reaching_the_thresholds_is_skipped(lib, mag):
    rng = np.random.default_rng(3 + mag)
    cells = np.concatenate([random_cells(rng, 4000, mag), tight_sound_cells(rng, 4000, mag)])
    _, exact, _ = screen(lib, cells, mag, 1.0, 0.0)
    lim = limits(exact, mag)
    um = exact[:, 0]
    checked = 0
    # thresholds placed on single cells' own values, to the last ulp, for
    # every limit in turn and for umach, from both sides
    for q in range(lim.shape[1]):
        col = lim[:, q]
        for idx in rng.choice(len(cells), 40, replace=False):
            base = col[idx] / K[7]
            for lp in (np.nextafter(base, 0.0), base, np.nextafter(base, np.inf)):
                missed, _, hit = screen(lib, cells, mag, lp, um.max())
                assert missed == 0, (q, idx, lp)
                if col[idx] <= lp * K[7]:
                    assert hit[idx], (q, idx)
                checked += 1
    for idx in rng.choice(len(cells), 60, replace=False):
        base = um[idx] / K[8]
        for up in (np.nextafter(base, 0.0), base, np.nextafter(base, np.inf)):
            missed, _, hit = screen(lib, cells, mag, lim.min() * 0.5, up)
            assert missed == 0, (idx, up)
            if um[idx] > up * K[8]:
                assert hit[idx], idx
            checked += 1
    assert checked >= 500


def test_unphysical_values_are_always_evaluated_and_rest_passes_umach(lib):
    rng = np.random.default_rng(11)
    cells = random_cells(rng, 8, False)
    bad = cells.copy()
    bad[0, 3] = -1.0
    bad[1, 4] = 0.0
    bad[2, 9] = -0.5
    bad[3, 0] = np.nan
    bad[4, 4] = np.nan
    bad[5, 3] = np.inf
    bad[6, 1] = np.inf
    _, _, hit = screen(lib, bad, False, 1e-30, 1e30)
    assert hit[:7].all(), hit
    assert not hit[7]
    rest = cells.copy()
    rest[:, 0:3] = 0.0
    missed, exact, hit = screen(lib, rest, False, 1e-30, 0.0)
    assert missed == 0 and not hit.any() and (exact[:, 0] == 0.0).all()


@pytest.mark.parametrize("mag", [False, True], ids=["hydro", "mhd"])
def test_proven_screened_extrema_are_the_exact_ones(lib, mag):
    rng = np.random.default_rng(21 + mag)
    cells = np.concatenate([random_cells(rng, 3000, mag), tight_sound_cells(rng, 1000, mag)])
    _, exact, _ = screen(lib, cells, mag, 1.0, 0.0)
    lmin, umax = limits(exact, mag).min(), exact[:, 0].max()
    kp = K.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    proven = refused = 0
    for f in np.concatenate([rng.uniform(0.3, 3.0, 150), [1.0, 1.0 / 1.05, 1.0 / 1.0500001]]):
        for g in (rng.uniform(0.3, 3.0), 1.0, 0.0):
            lp, up = lmin * f, umax * g
            _, ex, hit = screen(lib, cells, mag, lp, up)
            s_lim = limits(ex[hit], mag).min() if hit.any() else 1e300
            s_um = ex[hit, 0].max() if hit.any() else 0.0
            if lib.screen_proven_c(kp, s_lim, s_um, lp, up):
                proven += 1
                assert s_lim == lmin and s_um == umax, (f, g)
            else:
                refused += 1
    assert proven > 50 and refused > 50, (proven, refused)


def numpy_screen(cells, mag, lim_prev, um_prev, k, sound_factor=(5.0, 1.25), margin=1e-12):
    ru, rv, rw, ro, tt = (cells[:, i] for i in range(5))
    dd, rk = cells[:, 8], cells[:, 9]
    m2 = ru * ru + rv * rv + rw * rw
    thr, tum = lim_prev * k[7] * (1 + margin), um_prev * k[8] * (1 - margin)
    g2 = GAMMA * tt * ro * ro
    if mag:
        g2 = g2 + (cells[:, 5:8] ** 2).sum(1) * k[6] * ro
    hit = k[1] * m2 > tum * tt * ro * ro
    hit |= 0.5 * dd * dd * k[2] * k[3] * ro <= thr * rk
    hit |= 0.375 * dd * dd * k[4] * ro <= thr
    hit |= dd * dd * ro * ro <= thr * thr * (sound_factor[0] * m2 + sound_factor[1] * g2)
    if mag:
        hit |= 0.5 * dd * dd * k[5] <= thr
    return hit


def test_adversarial_cells_catch_plausible_screen_mistakes(lib):
    """The same data, screened by numpy restatements of the screen: the
    correct one misses nothing (as the C screen), while a sound bound of
    UU + G (dropping the 2 sqrt(UU G) term) or a margin below the rounding
    of the exact limits misses cells that reach the thresholds.  Re and
    Re Pr are large here, so that the sound-speed limit is the least one."""
    rng = np.random.default_rng(5)
    k = K.copy()
    k[2] = k[4] = 1e6
    cells = tight_sound_cells(rng, 4000, False)
    _, exact, _ = screen(lib, cells, False, 1.0, 0.0, k)
    snd = exact[:, 1]
    found = {"good": 0, "no_cross_term": 0, "no_margin": 0}
    for idx in rng.choice(len(cells), 200, replace=False):
        lp = np.nextafter(snd[idx] / k[7], np.inf)
        reach = limits(exact, False).min(1) <= lp * k[7]
        # um_prev far above every cell's ogamma*UU/T (all 1/16 here), so that
        # only the limit tests can select a cell
        missed, _, _ = screen(lib, cells, False, lp, 1e30, k)
        assert missed == 0
        for name, kw in (("good", {}), ("no_cross_term", {"sound_factor": (1.0, 1.0)}),
                         ("no_margin", {"margin": -1e-15})):
            h = numpy_screen(cells, False, lp, 1e30, k, **kw)
            found[name] += int((reach & ~h).sum())
    assert found["good"] == 0
    assert found["no_cross_term"] > 0 and found["no_margin"] > 0, found
