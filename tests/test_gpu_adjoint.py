"""Discrete adjoint (include/r3d_adj.h) against central finite differences of
the forward time step: dJ/dX0 along directions of every primary, dJ/dRe,
dJ/dtheta, and the torch.autograd.Function over the temperature perturbation.
Relative agreement required: 1e-6.

Two properties of J set how the differences are taken.
* Rounding noise: J (about 2e4 here) carries a rounding noise of a few parts in
  1e15 that changes with every perturbation of X0, so a central difference
  with step eps has an absolute error of roughly 1e-11 / eps on top of the
  eps^2 truncation error.  The direction is weighted towards the adjoint
  gradient, where the directional derivative is largest.  For the velocity,
  density and temperature the signal is then 1e2 .. 1e6 and eps = 1e-6 (of
  the field's magnitude) leaves errors near 1e-10; a sweep over eps shows the
  error falling as eps^2 down to 1e-6.  J is a thousand times less sensitive
  to the magnetic field and to Re: below eps ~ 3e-4 the noise dominates (error
  growing as 1 / eps), so those steps are 1e-3 (field) and 1e-4 (Re, theta).
* The CFL step dt = sff * min(limits) is smooth only while one cell and limit
  attain the minimum.  With the default Re = 2, Pr = 0.05 the thermal limit
  (proportional to rho) binds on the top plane, where rho is uniform to a few
  digits, and a step in rho moves the minimum between near-equal cells.  The
  CFL checks therefore run at Re = 20, Pr = 1, where the advective limit (set
  by the hottest cell) binds; the default physics is checked with a fixed dt
  (r3d_adj_set_fixed_dt), which makes J smooth everywhere.
"""
import importlib
import os
import sys

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

ARGS = ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "1", "--npez", "2"]
SMOOTH = ["--re", "20", "--pr", "1"]  # the advective dt limit binds
NSTEPS, EVERY = 6, 4
# finite-difference steps, relative to the field's magnitude or the parameter
EPS = {"ru": 1e-6, "rv": 1e-6, "rw": 1e-6, "ro": 1e-6, "tt": 1e-6,
       "bx": 1e-3, "by": 1e-3, "bz": 1e-3, "re": 1e-4, "theta": 1e-4, "tb": 1e-4}
PRIM = ["ru", "rv", "rw", "ro", "tt"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    """A developed state (10 plain steps from the initial condition) as X0."""
    run = pkg.Run(ARGS + SMOOTH)
    sim = run.make_sim()
    sim.step(10)
    x0 = sim.download_global(PRIM)
    sim.close()
    return run, x0


@pytest.fixture(scope="module")
def fixed_setup():
    """Default physics, a developed state, and 0.8 of its CFL step as the
    fixed dt."""
    run = pkg.Run(ARGS)
    sim = run.make_sim()
    sim.step(10)
    x0 = sim.download_global(PRIM)
    dt = 0.8 * sim.scalars()["dt"]
    sim.close()
    return run, x0, dt


def fd_check(cfg, met, x0, field, grads, dre, dth, fixed_dt=0.0, tol=1e-6, dtb=None):
    """Central difference of J along a direction of one primary, or in Re /
    theta / the bottom temperature tb, against the adjoint's directional
    derivative.  The direction is
    the adjoint gradient scaled to unit maximum plus uniform noise in [-1, 1],
    times the field's largest magnitude; the step is EPS[field] of that."""
    eps = EPS[field]
    if field in ("re", "theta", "tb"):
        v0 = getattr(cfg, field)
        h = eps * v0
        fd = (run_J(pkg.cfg_with(cfg, **{field: v0 + h}), met, x0, fixed_dt=fixed_dt) -
              run_J(pkg.cfg_with(cfg, **{field: v0 - h}), met, x0, fixed_dt=fixed_dt)) / (2 * h)
        ad = {"re": dre, "theta": dth, "tb": dtb}[field]
    else:
        rng = np.random.default_rng(sum(map(ord, field)))
        g = grads[field][1:-1, 1:-1, 1:-1]
        d = np.zeros_like(x0[field])
        d[1:-1, 1:-1, 1:-1] = g / np.abs(g).max() + rng.uniform(-1, 1, g.shape)
        d *= np.abs(x0[field][1:-1, 1:-1, 1:-1]).max()
        xp = {k: v.copy() for k, v in x0.items()}
        xm = {k: v.copy() for k, v in x0.items()}
        xp[field] += eps * d
        xm[field] -= eps * d
        fd = (run_J(cfg, met, xp, fixed_dt=fixed_dt) -
              run_J(cfg, met, xm, fixed_dt=fixed_dt)) / (2 * eps)
        ad = float(np.sum(grads[field] * d))
    say(f"[gpu]   adjoint {ad:.12e} fd {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) <= tol * abs(fd), (field, ad, fd)


def run_J(cfg, metrics, x0, nsteps=NSTEPS, fixed_dt=0.0):
    adj = pkg.Adjoint(cfg, metrics, EVERY)
    adj.set_fixed_dt(fixed_dt)
    adj.set_initial(x0)
    J = adj.forward(nsteps)
    adj.close()
    return J


@pytest.mark.parametrize("field", PRIM + ["re", "theta", "tb"])
def test_fixed_dt_gradient_matches_finite_differences(fixed_setup, field):
    """With a fixed step J is smooth: the transpose of the fluxes and the
    boundary conditions alone, at the default Re = 2, Pr = 0.05."""
    run, x0, dt = fixed_setup
    cfg, met = run.cfg(), run.metrics()
    say(f"[gpu] fixed-dt adjoint dJ/d{field} (dt {dt:.6e})")
    adj = pkg.Adjoint(cfg, met, EVERY)
    adj.set_fixed_dt(dt)
    adj.set_initial(x0)
    adj.forward(NSTEPS)
    grads, dre, dth = adj.backward()
    dtb = adj.dJ_dtb()
    adj.close()
    fd_check(cfg, met, x0, field, grads, dre, dth, fixed_dt=dt, dtb=dtb)


def test_adjoint_forward_is_the_plain_path(setup):
    run, x0 = setup
    say("[gpu] adjoint forward vs plain path")
    adj = pkg.Adjoint(run.cfg(), run.metrics(), EVERY)
    adj.set_initial(x0)
    adj.forward(NSTEPS)
    got = adj.final_state()
    sim = pkg.HipSim(run.cfg())
    sim.set_metrics(run.metrics())
    sim.upload_global(x0)
    sim.set_scalars()
    sim.step(NSTEPS)
    want = sim.download_global(PRIM)
    for n in PRIM:
        a, b = got[n][1:-1, 1:-1, 1:-1], want[n][1:-1, 1:-1, 1:-1]
        assert np.array_equal(a, b), n


@pytest.mark.parametrize("field", PRIM + ["re", "theta"])
def test_cfl_gradient_matches_finite_differences(setup, field):
    run, x0 = setup
    cfg, met = run.cfg(), run.metrics()
    say(f"[gpu] adjoint dJ/d{field}")
    adj = pkg.Adjoint(cfg, met, EVERY)
    adj.set_initial(x0)
    adj.forward(NSTEPS)
    grads, dre, dth = adj.backward()
    fwd_ms, bwd_ms = adj.timing()
    adj.close()
    say(f"[gpu]   forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms")
    fd_check(cfg, met, x0, field, grads, dre, dth)


def test_torch_function_gradient_of_the_perturbation():
    """J(delta) with tt0 = T_static (1 + delta) through torch autograd, from the
    reference's initial condition, against finite differences."""
    torch = pytest.importorskip("torch")
    at = importlib.import_module("rast3dmhd_amd.adjoint_torch")
    say("[gpu] torch.autograd adjoint")
    prob = at.KineticEnergyRun(ARGS, NSTEPS, EVERY)
    static = pkg.Run(ARGS + ["--ampt", "0"]).state()["tt"][1:-1, 1:-1, 1:-1]
    t_static = torch.from_numpy(static.copy())
    delta0 = prob.initial_tt() / t_static - 1.0
    delta = delta0.clone().requires_grad_(True)
    re = torch.tensor(prob.cfg.re, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(prob.cfg.theta, dtype=torch.float64, requires_grad=True)
    J = prob(t_static * (1.0 + delta), re, th)
    J.backward()
    rng = np.random.default_rng(3)
    d = torch.from_numpy(rng.uniform(-1, 1, tuple(delta.shape)) * 0.005)
    eps = 1e-6
    with torch.no_grad():
        Jp = prob(t_static * (1.0 + delta0 + eps * d), re, th)
        Jm = prob(t_static * (1.0 + delta0 - eps * d), re, th)
    fd = float((Jp - Jm) / (2 * eps))
    ad = float(torch.sum(delta.grad * d))
    say(f"[gpu]   autograd {ad:.12e} fd {fd:.12e}; dJ/dRe {float(re.grad):.6e}")
    assert abs(ad - fd) <= 1e-6 * abs(fd)
    assert re.grad is not None and th.grad is not None


def test_torch_function_on_gpu_tensors_equals_host_tensors():
    """The torch binding with GPU tensors (state and gradients in device
    memory, r3d_adj_set_initial_device / r3d_adj_backward_device) gives J
    bit for bit as with CPU tensors (host arrays), and the gradients to the
    run-to-run reproducibility of the transpose, whose atomic sums have no
    fixed order (1e-12 of each field's largest entry)."""
    torch = pytest.importorskip("torch")
    at = importlib.import_module("rast3dmhd_amd.adjoint_torch")
    say("[gpu] torch adjoint: GPU tensors against CPU tensors")
    prob = at.KineticEnergyRun(ARGS, NSTEPS, EVERY)
    out = []
    for dev in (torch.device("cpu"), torch.device("cuda:0")):
        tt0 = prob.initial_tt().to(dev).requires_grad_(True)
        re = torch.tensor(prob.cfg.re, dtype=torch.float64, device=dev, requires_grad=True)
        th = torch.tensor(prob.cfg.theta, dtype=torch.float64, device=dev, requires_grad=True)
        J = prob(tt0, re, th)
        J.backward()
        assert J.device.type == dev.type and tt0.grad.device.type == dev.type
        out.append((float(J.detach()), tt0.grad.cpu().numpy(), float(re.grad), float(th.grad)))
    (ja, ga, ra, ta), (jb, gb, rb, tb) = out

    def close(a, b):
        return np.abs(np.asarray(a) - np.asarray(b)).max() <= 1e-12 * np.abs(a).max()

    assert ja == jb
    assert close(ga, gb) and close(ra, rb) and close(ta, tb)
    # the state form, B and the density included
    st = prob.static_state(torch.tensor(prob.cfg.theta, dtype=torch.float64))
    res = []
    for dev in (torch.device("cpu"), torch.device("cuda:0")):
        f = {n: st[n].detach().to(dev).clone().requires_grad_(True)
             for n in ("ro", "tt", "bx", "by", "bz")}
        J = prob.kinetic_energy(f, torch.tensor(prob.cfg.re, dtype=torch.float64, device=dev),
                                torch.tensor(prob.cfg.theta, dtype=torch.float64, device=dev),
                                torch.tensor(prob.cfg.tb, dtype=torch.float64, device=dev))
        J.backward()
        res.append((float(J.detach()), f["ro"].grad.cpu().numpy(), f["tt"].grad.cpu().numpy()))
    assert res[0][0] == res[1][0]
    for a, b in zip(res[0][1:], res[1][1:]):
        assert close(a, b)
    prob.close()


MAG_ARGS = ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2",
            "--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2",
            "--bzp", "10", "--lrot", "1", "--ry", "0.1", "--ang", "0"] + SMOOTH
MPRIM = PRIM + ["bx", "by", "bz"]


@pytest.fixture(scope="module")
def msetup():
    """A developed MHD state with a three-dimensional field as X0."""
    run = pkg.Run(MAG_ARGS)
    sim = run.make_sim()
    sim.step(6)
    x0 = sim.download_global(MPRIM)
    sim.close()
    rng = np.random.default_rng(4)
    for n in ("bx", "by", "bz"):
        x0[n][1:-1, 1:-1, 1:-1] += 0.05 * rng.uniform(-1, 1, x0[n][1:-1, 1:-1, 1:-1].shape)
    return run, x0


@pytest.mark.parametrize("field", MPRIM + ["re", "theta"])
def test_mhd_gradient_matches_finite_differences(msetup, field):
    run, x0 = msetup
    cfg, met = run.cfg(), run.metrics()
    say(f"[gpu] MHD adjoint dJ/d{field}")
    adj = pkg.Adjoint(cfg, met, EVERY)
    adj.set_initial(x0)
    adj.forward(NSTEPS)
    grads, dre, dth = adj.backward()
    adj.close()
    fd_check(cfg, met, x0, field, grads, dre, dth)


@pytest.mark.parametrize("field", MPRIM + ["re", "theta"])
def test_mhd_fixed_dt_gradient_matches_finite_differences(msetup, field):
    run, x0 = msetup
    cfg, met = run.cfg(), run.metrics()
    sim = pkg.HipSim(cfg)
    sim.set_metrics(met)
    sim.upload_global(x0)
    sim.set_scalars()
    sim.step(1)
    dt = 0.8 * sim.scalars()["dt"]
    sim.close()
    say(f"[gpu] MHD fixed-dt adjoint dJ/d{field} (dt {dt:.6e})")
    adj = pkg.Adjoint(cfg, met, EVERY)
    adj.set_fixed_dt(dt)
    adj.set_initial(x0)
    adj.forward(NSTEPS)
    grads, dre, dth = adj.backward()
    adj.close()
    fd_check(cfg, met, x0, field, grads, dre, dth, fixed_dt=dt)


@pytest.mark.parametrize("mag", [False, True], ids=["hydro", "mhd"])
def test_tiled_transpose_matches_per_cell_scatter(setup, msetup, mag):
    """The LDS-tiled flux transpose (default) against the per-cell kernel with
    global atomics (R3D_ADJ_VJP=atomic), whose gradients the finite-difference
    checks above pin: the same sums in another order."""
    run, x0 = msetup if mag else setup
    say(f"[gpu] tiled vs per-cell transpose ({'mhd' if mag else 'hydro'})")
    out = []
    for mode in ("atomic", "tiled"):
        os.environ["R3D_ADJ_VJP"] = mode
        try:
            adj = pkg.Adjoint(run.cfg(), run.metrics(), EVERY)
        finally:
            del os.environ["R3D_ADJ_VJP"]
        adj.set_initial(x0)
        adj.forward(NSTEPS)
        out.append(adj.backward())
        adj.close()
    (ga, rea, tha), (gt, ret, tht) = out
    for n in ga:
        scale = np.abs(ga[n]).max()
        err = np.abs(ga[n] - gt[n]).max()
        say(f"[gpu]   {n}: max |diff| {err:.3e} of {scale:.3e}")
        assert err <= 1e-11 * scale, n
    assert abs(rea - ret) <= 1e-11 * abs(rea) and abs(tha - tht) <= 1e-11 * abs(tha)


@pytest.mark.parametrize("mag", [False, True], ids=["hydro", "mhd"])
@pytest.mark.parametrize("nsteps", [NSTEPS, EVERY - 1], ids=["two_segments", "one_segment"])
def test_repeated_backward_gives_the_same_gradient(setup, msetup, mag, nsteps):
    """The forward leaves its last segment's stage states and dt in the
    segment buffers, and the first backward transposes that segment without
    a recompute.  A backward that recomputes an earlier segment into those
    buffers must recompute the last one the next time: two backward calls
    after one forward give the same gradient (to the sum-order rounding of
    the transposes' atomics), with one segment and with two."""
    run, x0 = msetup if mag else setup
    say(f"[gpu] repeated backward ({'mhd' if mag else 'hydro'}, {nsteps} steps, every {EVERY})")
    adj = pkg.Adjoint(run.cfg(), run.metrics(), EVERY)
    adj.set_initial(x0)
    adj.forward(nsteps)
    (g1, re1, th1), (g2, re2, th2) = adj.backward(), adj.backward()
    adj.close()
    for n in g1:
        scale = np.abs(g1[n]).max()
        err = np.abs(g1[n] - g2[n]).max()
        say(f"[gpu]   {n}: max |diff| {err:.3e} of {scale:.3e}")
        assert err <= 1e-11 * scale, n
    assert abs(re1 - re2) <= 1e-11 * abs(re1) and abs(th1 - th2) <= 1e-11 * abs(th1)


RAGGED = [
    (["--npx", "20", "--npy", "12", "--npz", "30", "--npey", "1", "--npez", "3"], False),
    (["--npx", "36", "--npy", "18", "--npz", "21", "--npey", "2", "--npez", "3", "--lrot", "1",
      "--ry", "0.2", "--ang", "0.3"], False),
    (["--npx", "18", "--npy", "20", "--npz", "26", "--npey", "2", "--npez", "2", "--lmag", "1",
      "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2", "--bzp", "10"], True),
    (["--npx", "12", "--npy", "33", "--npz", "24", "--npey", "1", "--npez", "1", "--lmag", "1",
      "--rm", "20", "--beta", "5", "--ampb", "1", "--bfh", "2", "--bzp", "10"], True),
]


@pytest.mark.parametrize("case", range(len(RAGGED)),
                         ids=["hydro_20x12x30", "rot_36x18x21", "mhd_18x20x26", "mhd_12x33x24_1z"])
def test_tiled_transpose_on_partial_tiles(case):
    """Grids that are no multiple of the transpose tile (16 x 16) in x or y,
    chunks of uneven height, one z rank: the tiled transpose against the
    per-cell kernel with global atomics."""
    args, mag = RAGGED[case]
    run = pkg.Run(args)
    names = MPRIM if mag else PRIM
    sim = run.make_sim()
    sim.step(4)
    x0 = sim.download_global(names)
    sim.close()
    if mag:
        rng = np.random.default_rng(9)
        for n in ("bx", "by", "bz"):
            x0[n][1:-1, 1:-1, 1:-1] += 0.05 * rng.uniform(-1, 1, x0[n][1:-1, 1:-1, 1:-1].shape)
    say(f"[gpu] tiled vs per-cell transpose, {' '.join(args[:6])}")
    out = []
    for mode in ("atomic", "tiled"):
        os.environ["R3D_ADJ_VJP"] = mode
        try:
            adj = pkg.Adjoint(run.cfg(), run.metrics(), 3)
        finally:
            del os.environ["R3D_ADJ_VJP"]
        adj.set_initial(x0)
        adj.forward(5)
        out.append(adj.backward())
        adj.close()
    (ga, rea, tha), (gt, ret, tht) = out
    for n in ga:
        scale = np.abs(ga[n]).max()
        err = np.abs(ga[n] - gt[n]).max()
        assert err <= 1e-11 * scale, (n, err, scale)
    assert abs(rea - ret) <= 1e-11 * abs(rea) and abs(tha - tht) <= 1e-11 * abs(tha)


@pytest.mark.parametrize("field", ["ru", "ro", "tt", "bx", "re"])
def test_fixed_dt_gradient_on_a_ragged_mhd_grid(field):
    """Finite differences on the ragged MHD grid (fixed dt): the partial-tile
    paths of the transpose carry the right entries, not merely the same ones
    as the per-cell kernel."""
    args, _ = RAGGED[2]
    run = pkg.Run(args)
    sim = run.make_sim()
    sim.step(4)
    x0 = sim.download_global(MPRIM)
    dt = 0.8 * sim.scalars()["dt"]
    sim.close()
    rng = np.random.default_rng(9)
    for n in ("bx", "by", "bz"):
        x0[n][1:-1, 1:-1, 1:-1] += 0.05 * rng.uniform(-1, 1, x0[n][1:-1, 1:-1, 1:-1].shape)
    cfg, met = run.cfg(), run.metrics()
    say(f"[gpu] ragged MHD fixed-dt adjoint dJ/d{field}")
    adj = pkg.Adjoint(cfg, met, EVERY)
    adj.set_fixed_dt(dt)
    adj.set_initial(x0)
    adj.forward(NSTEPS)
    grads, dre, dth = adj.backward()
    adj.close()
    fd_check(cfg, met, x0, field, grads, dre, dth, fixed_dt=dt)


@pytest.mark.parametrize("mag", [False, True], ids=["hydro", "mhd"])
def test_total_theta_derivative_through_the_static_profile(mag):
    """theta sets the static profile (rho, T and the bottom temperature tb) as
    well as the heat-flux conditions.  J(theta) = J of the profile built at
    theta with the reference's temperature perturbation on top: the autograd
    total (adjoint for the run, central difference of the host setup for the
    profile, dJ/dtb from the transposed boundary pass) against central
    differences of the whole pipeline.  Fixed dt, so J is smooth in theta."""
    torch = pytest.importorskip("torch")
    at = importlib.import_module("rast3dmhd_amd.adjoint_torch")
    args = MAG_ARGS if mag else ARGS
    say(f"[gpu] total dJ/dtheta through the static profile ({'mhd' if mag else 'hydro'})")
    sim = pkg.Run(args).make_sim()
    sim.step(1)
    dt = 0.8 * sim.scalars()["dt"]
    sim.close()
    prob = at.KineticEnergyRun(args, NSTEPS, EVERY, fixed_dt=dt)
    th0 = prob.cfg.theta
    t_static = pkg.Run(args + ["--ampt", "0"]).state()["tt"][1:-1, 1:-1, 1:-1]
    delta = prob.initial_tt() / torch.from_numpy(t_static.copy()) - 1.0
    re = torch.tensor(prob.cfg.re, dtype=torch.float64)

    def J_of(theta):
        st = prob.static_state(theta)
        fields = {n: st[n] for n in ("ro", "bx", "by", "bz")}
        fields["tt"] = st["tt"] * (1.0 + delta)
        return prob.kinetic_energy(fields, re, theta, st["tb"])

    th = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    J = J_of(th)
    J.backward()
    ad = float(th.grad)
    h = EPS["theta"] * th0
    with torch.no_grad():
        fd = float(J_of(torch.tensor(th0 + h, dtype=torch.float64)) -
                   J_of(torch.tensor(th0 - h, dtype=torch.float64))) / (2 * h)
    with torch.no_grad():
        st = prob.static_state(torch.tensor(th0, dtype=torch.float64))
    say(f"[gpu]   J {float(J.detach()):.12e} autograd {ad:.12e} fd {fd:.12e}")
    assert abs(ad - fd) <= 1e-6 * abs(fd)
    assert abs(float(st["tb"]) - prob.cfg.tb) <= 1e-12 * abs(prob.cfg.tb)
