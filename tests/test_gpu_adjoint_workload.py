"""The adjoint sensitivity workload: rotating MHD with a magnetic layer at
128 x 128 x 256 (virtual 2 x 4 ranks), 200 RK3 steps, checkpoints every 20
steps, J = volume-integrated kinetic energy at the final time.  The gradient
with respect to the initial temperature perturbation delta (tt0 = T_static
(1 + delta)), Re and theta comes from the torch.autograd.Function over the
hand-written adjoint and is checked against central finite differences of
the forward run:

* 16 seeded random directions d of delta, uniform in [-1, 1] per cell,
  delta +- 1e-6 d; agreement 1e-6 relative (measured: at most 1.1e-7);
* Re +- 1e-6 Re (measured 6e-10);
* theta +- 1e-5 theta (measured 7e-7).  J is about 1.1e5 and dJ/dtheta
  about 6, so the difference quotient of a 1e-6 step sits on J's rounding
  noise of a few parts in 1e15 (error 2e-5); 1e-5 is the smallest step
  clear of it.
"""
import importlib
import sys

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

MHD = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2", "--bzp", "10",
       "--lrot", "1", "--ry", "0.1", "--ang", "0"]
ARGS = ["--npx", "128", "--npy", "128", "--npz", "256", "--npey", "2", "--npez", "4"] + MHD
NSTEPS, EVERY = 200, 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def workload():
    torch = pytest.importorskip("torch")
    at = importlib.import_module("rast3dmhd_amd.adjoint_torch")
    say("[gpu] adjoint workload: MHD 128x128x256, 200 steps")
    prob = at.KineticEnergyRun(ARGS, NSTEPS, EVERY)
    static = pkg.Run(ARGS + ["--ampt", "0"]).state()["tt"][1:-1, 1:-1, 1:-1]
    # GPU tensors: the state and the gradients stay in device memory
    dev = torch.device("cuda:0")
    ts = torch.from_numpy(static.copy()).to(dev)
    delta0 = prob.initial_tt().to(dev) / ts - 1.0
    delta = delta0.clone().requires_grad_(True)
    re = torch.tensor(prob.cfg.re, dtype=torch.float64, device=dev, requires_grad=True)
    th = torch.tensor(prob.cfg.theta, dtype=torch.float64, device=dev, requires_grad=True)
    J = prob(ts * (1.0 + delta), re, th)
    J.backward()
    say(f"[gpu]   J = {float(J.detach()):.10e}, dJ/dRe = {float(re.grad):.10e}, "
        f"dJ/dtheta = {float(th.grad):.10e}")

    def J_at(dl, r, t):
        with torch.no_grad():
            return float(prob(ts * (1.0 + dl), torch.tensor(r, dtype=torch.float64, device=dev),
                              torch.tensor(t, dtype=torch.float64, device=dev)))

    return torch, prob, delta0, delta.grad.clone(), float(re.grad), float(th.grad), J_at


def test_perturbation_gradient_in_16_random_directions(workload):
    torch, prob, delta0, g, _, _, J_at = workload
    rng = np.random.default_rng(2024)
    eps, worst = 1e-6, 0.0
    re, th = prob.cfg.re, prob.cfg.theta
    for n in range(16):
        d = torch.from_numpy(rng.uniform(-1, 1, tuple(delta0.shape))).to(delta0.device)
        ad = float(torch.sum(g * d))
        fd = (J_at(delta0 + eps * d, re, th) - J_at(delta0 - eps * d, re, th)) / (2 * eps)
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        say(f"[gpu]   direction {n}: adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e}")
        assert rel <= 1e-6, (n, ad, fd)
    say(f"[gpu]   worst {worst:.2e}")


@pytest.mark.parametrize("param,step", [("re", 1e-6), ("theta", 1e-5)])
def test_parameter_gradient(workload, param, step):
    _, prob, delta0, _, dre, dth, J_at = workload
    re, th = prob.cfg.re, prob.cfg.theta
    if param == "re":
        h = step * re
        fd = (J_at(delta0, re + h, th) - J_at(delta0, re - h, th)) / (2 * h)
        ad = dre
    else:
        h = step * th
        fd = (J_at(delta0, re, th + h) - J_at(delta0, re, th - h)) / (2 * h)
        ad = dth
    rel = abs(ad - fd) / abs(fd)
    say(f"[gpu]   dJ/d{param}: adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e}")
    assert rel <= 1e-6, (param, ad, fd)


def test_total_theta_derivative_at_workload_size(workload):
    """theta sets the static profile (rho, T, B and the fixed bottom
    temperature tb) as well as the heat-flux conditions.  With the workload's
    perturbation delta held fixed on top of the profile, the autograd total
    dJ/dtheta (adjoint for the 200-step run with its CFL dt, central
    difference of the host setup for the profile, dJ/dtb from the transposed
    boundary pass) against central differences of the whole pipeline, the
    profile rebuilt at theta +- 1e-5 theta."""
    torch, prob, delta0, _, _, dth_bc, _ = workload
    dev = delta0.device
    delta = delta0.detach().cpu()
    re = torch.tensor(prob.cfg.re, dtype=torch.float64)
    th0 = prob.cfg.theta

    def J_of(theta):
        st = prob.static_state(theta)
        fields = {n: st[n] for n in ("ro", "bx", "by", "bz")}
        fields["tt"] = st["tt"] * (1.0 + delta)
        return prob.kinetic_energy(fields, re, theta, st["tb"])

    th = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    J = J_of(th)
    J.backward()
    ad = float(th.grad)
    h = 1e-5 * th0
    with torch.no_grad():
        fd = float(J_of(torch.tensor(th0 + h, dtype=torch.float64)) -
                   J_of(torch.tensor(th0 - h, dtype=torch.float64))) / (2 * h)
    rel = abs(ad - fd) / abs(fd)
    say(f"[gpu]   total dJ/dtheta: adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e} "
        f"(boundary part alone {dth_bc:.10e}); J {float(J.detach()):.10e} on {dev}")
    assert rel <= 1e-6, (ad, fd)
