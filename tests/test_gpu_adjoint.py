"""Discrete adjoint (include/r3d_adj.h) against central finite differences of
the forward time step: dJ/dX0 in random directions of every primary, dJ/dRe,
dJ/dtheta, and the torch.autograd.Function over the temperature perturbation.
Relative agreement required: 1e-6.

J depends on the state through dt = sff * min(limits), which is smooth only
while one cell and limit attain the minimum.  With the default Re = 2,
Pr = 0.05 the thermal limit (proportional to rho) binds on the top plane,
where rho is uniform to a few digits: a finite-difference step in rho then
moves the minimum between near-equal cells and the difference quotient no
longer measures a derivative.  The state checks therefore run at Re = 20,
Pr = 1, where the advective limit (set by the hottest cell) binds; the
default physics is covered by the temperature-perturbation check, whose
directions leave rho, and with it the binding limit, unchanged.
"""
import importlib
import sys

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

ARGS = ["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "1", "--npez", "2"]
SMOOTH = ["--re", "20", "--pr", "1"]  # the advective dt limit binds
NSTEPS, EVERY = 6, 4
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


def run_J(cfg, metrics, x0, nsteps=NSTEPS):
    adj = pkg.Adjoint(cfg, metrics, EVERY)
    adj.set_initial(x0)
    J = adj.forward(nsteps)
    adj.close()
    return J


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


@pytest.mark.parametrize("field", PRIM)
def test_state_gradient_matches_finite_differences(setup, field):
    run, x0 = setup
    cfg, met = run.cfg(), run.metrics()
    say(f"[gpu] adjoint dJ/d{field}")
    adj = pkg.Adjoint(cfg, met, EVERY)
    adj.set_initial(x0)
    adj.forward(NSTEPS)
    grads, _, _ = adj.backward()
    fwd_ms, bwd_ms = adj.timing()
    adj.close()
    rng = np.random.default_rng(hash(field) % 1000)
    d = np.zeros_like(x0[field])
    d[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, d[1:-1, 1:-1, 1:-1].shape)
    d *= np.abs(x0[field][1:-1, 1:-1, 1:-1]).max()
    eps = 1e-6
    xp = {k: v.copy() for k, v in x0.items()}
    xm = {k: v.copy() for k, v in x0.items()}
    xp[field] += eps * d
    xm[field] -= eps * d
    fd = (run_J(cfg, met, xp) - run_J(cfg, met, xm)) / (2 * eps)
    ad = float(np.sum(grads[field] * d))
    say(f"[gpu]   adjoint {ad:.12e} fd {fd:.12e} (forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms)")
    assert abs(ad - fd) <= 1e-6 * abs(fd), (field, ad, fd)


@pytest.mark.parametrize("param", ["re", "theta"])
def test_parameter_gradients_match_finite_differences(setup, param):
    run, x0 = setup
    cfg, met = run.cfg(), run.metrics()
    say(f"[gpu] adjoint dJ/d{param}")
    adj = pkg.Adjoint(cfg, met, EVERY)
    adj.set_initial(x0)
    adj.forward(NSTEPS)
    _, dre, dth = adj.backward()
    adj.close()
    v0 = cfg.re if param == "re" else cfg.theta
    h = 1e-6 * v0
    Jp = run_J(pkg.cfg_with(cfg, **{param: v0 + h}), met, x0)
    Jm = run_J(pkg.cfg_with(cfg, **{param: v0 - h}), met, x0)
    fd = (Jp - Jm) / (2 * h)
    ad = dre if param == "re" else dth
    say(f"[gpu]   adjoint {ad:.12e} fd {fd:.12e}")
    assert abs(ad - fd) <= 1e-6 * abs(fd), (param, ad, fd)


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
    if field in ("re", "theta"):
        v0 = cfg.re if field == "re" else cfg.theta
        h = 1e-6 * v0
        fd = (run_J(pkg.cfg_with(cfg, **{field: v0 + h}), met, x0) -
              run_J(pkg.cfg_with(cfg, **{field: v0 - h}), met, x0)) / (2 * h)
        ad = dre if field == "re" else dth
    else:
        rng = np.random.default_rng(len(field))
        d = np.zeros_like(x0[field])
        d[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, d[1:-1, 1:-1, 1:-1].shape)
        d *= np.abs(x0[field][1:-1, 1:-1, 1:-1]).max()
        eps = 1e-6
        xp = {k: v.copy() for k, v in x0.items()}
        xm = {k: v.copy() for k, v in x0.items()}
        xp[field] += eps * d
        xm[field] -= eps * d
        fd = (run_J(cfg, met, xp) - run_J(cfg, met, xm)) / (2 * eps)
        ad = float(np.sum(grads[field] * d))
    say(f"[gpu]   adjoint {ad:.12e} fd {fd:.12e}")
    assert abs(ad - fd) <= 1e-6 * abs(fd), (field, ad, fd)
