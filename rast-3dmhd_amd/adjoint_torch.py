"""torch.autograd binding of the discrete adjoint (include/r3d_adj.h).

KineticEnergyRun(args, nsteps) holds a run built by the host setup from the
reference CLI flags `args` (hydrodynamic or MHD) and offers two
differentiable maps to J = the final kinetic energy:

* prob(tt0, Re, theta): tt0 is the interior initial temperature
  (npz, npy, npx) replacing that of the run's initial state; theta enters
  through the heat-flux boundary conditions only (the density, the field and
  the bottom temperature stay those of the run's setup).  Composing
  tt0 = T_static * (1 + delta) in torch gives the gradient with respect to
  the perturbation field delta.
* prob.static_state(theta) and prob.kinetic_energy(fields, Re, theta, tb):
  the static profile (rho, T, B and the bottom temperature tb) as a function
  of theta, and J of any initial state.  Composed as
      st = prob.static_state(theta)
      J = prob.kinetic_energy({"ro": st["ro"], "tt": st["tt"] * (1 + delta), ...},
                              re, theta, st["tb"])
  they give the total derivative with respect to theta, through the initial
  profile as well as the boundary conditions.

The forward is the plain bit-exact time step on the GPU; the backward is the
hand-written adjoint.  With GPU tensors the initial state and the gradients
stay in device memory (r3d_adj_set_initial_device / r3d_adj_backward_device:
the setup's state is copied to the device once per problem, and each call
only writes the tensors into a device copy of it); CPU tensors go through
host arrays.  The static profile's theta derivative is the central
difference of the host setup (a smooth ODE solution; step 1e-6 theta).
theta also sets the conductivity profile when the conductivity jump is on
(--pzp != 0); that dependence is not differentiated (the metrics are held
fixed), so the total derivative is complete for --pzp 0 only.
"""
import numpy as np
import torch

from . import Adjoint, Run, cfg_with


class KineticEnergyRun:
    def __init__(self, args, nsteps, checkpoint_every=20, fixed_dt=0.0):
        self.args = list(args)
        self.fixed_dt = float(fixed_dt)  # > 0: every step takes this dt
        self.run = Run(args)
        self.base = dict(self.run.state())
        self.metrics = self.run.metrics()
        self.cfg = self.run.cfg()
        self.nsteps = nsteps
        self.every = checkpoint_every
        self._key = None
        self._adj = None
        self._base_dev = None  # the setup's state as device tensors, made on first use

    def base_on(self, device):
        """A fresh copy of the setup's full initial state (ghosts included) as
        float64 tensors on `device` (the device copy of the host state is made
        once and cloned per call)."""
        if device.type == "cpu":
            return {k: torch.from_numpy(v.copy()) for k, v in self.base.items()}
        if self._base_dev is None or self._base_dev["tt"].device != device:
            self._base_dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to(device)
                              for k, v in self.base.items()}
        return {k: v.clone() for k, v in self._base_dev.items()}

    def close(self):
        """Free the cached adjoint context (device checkpoints and segment
        buffers) and the host setup."""
        if self._adj is not None:
            self._adj.close()
            self._adj, self._key = None, None
        self._base_dev = None
        self.run.close()

    def adjoint(self, re, theta):
        """The one adjoint context this problem keeps on the device: reused
        while (Re, theta) stay the same, closed and replaced when they change."""
        key = (re, theta)
        if key != self._key:
            if self._adj is not None:
                self._adj.close()
            self._adj = Adjoint(cfg_with(self.cfg, re=re, theta=theta), self.metrics, self.every)
            self._adj.set_fixed_dt(self.fixed_dt)
            self._key = key
        return self._adj

    def initial_tt(self):
        """The run's own interior initial temperature."""
        return torch.from_numpy(self.base["tt"][1:-1, 1:-1, 1:-1].copy())

    def __call__(self, tt0, re, theta):
        return _KineticEnergy.apply(tt0, re, theta, self)

    def static_state(self, theta):
        """Interior static fields of the setup at this theta (no perturbation)
        as tensors differentiable in theta: "ro", "tt", "bx", "by", "bz" (zero
        without a field) and the bottom temperature "tb" (0-d)."""
        if np.any(np.asarray(self.metrics["rkapa"]) != 1.0):
            raise NotImplementedError(
                "static_state: the conductivity profile (--pzp != 0) depends on theta and "
                "is not differentiated")
        out = _StaticState.apply(theta, self)
        return dict(zip(("ro", "tt", "bx", "by", "bz", "tb"), out))

    def static_numpy(self, theta):
        run = Run(self.args + ["--theta", repr(float(theta)), "--ampt", "0"])
        st = run.state()
        sl = (slice(1, -1),) * 3
        zero = np.zeros_like(st["ro"][sl])
        arrs = [st["ro"][sl].copy(), st["tt"][sl].copy()] + [
            st[n][sl].copy() if self.cfg.lmag else zero for n in ("bx", "by", "bz")]
        return arrs, float(run.cfg().tb)

    def kinetic_energy(self, fields, re, theta, tb):
        """J of the initial state whose interior density, temperature and (with
        LMAG) field are the tensors fields["ro"], fields["tt"], fields["bx"..
        "bz"] (the velocities are the setup's), with Re, theta (heat-flux
        conditions) and the fixed bottom temperature tb."""
        zero = torch.zeros_like(fields["tt"])
        b = [fields.get(n, zero) for n in ("bx", "by", "bz")]
        return _KineticEnergyState.apply(fields["ro"], fields["tt"], *b, re, theta, tb, self)


def _set_initial(adj, fields):
    """X0 from full-state tensors: device pointers for GPU tensors (the
    writes of torch's stream completed first), host arrays for CPU ones."""
    if fields["tt"].is_cuda:
        torch.cuda.synchronize(fields["tt"].device)
        adj.set_initial_device({n: fields[n].data_ptr() for n in adj.names})
    else:
        adj.set_initial({n: fields[n].numpy() for n in adj.names})


def _devices(*xs):
    """The device of each tensor argument (None for plain numbers), so that
    every gradient is returned where its input lives."""
    return [x.device if torch.is_tensor(x) else None for x in xs]


def _scalar(v, device):
    return torch.tensor(v, dtype=torch.float64, device=device)


def _backward(adj, names, device):
    """Gradients of the wanted primaries as full-state tensors on `device`,
    and dJ/dRe, dJ/dtheta."""
    if device.type != "cpu":
        out = {n: torch.zeros(adj.shape, dtype=torch.float64, device=device) for n in names}
        torch.cuda.synchronize(device)
        dre, dth = adj.backward_device({n: out[n].data_ptr() for n in names})
        return out, dre, dth
    grads, dre, dth = adj.backward()
    return {n: torch.from_numpy(grads[n]) for n in names}, dre, dth


class _KineticEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tt0, re, theta, prob):
        adj = prob.adjoint(float(re), float(theta))
        fields = prob.base_on(tt0.device)
        fields["tt"][1:-1, 1:-1, 1:-1] = tt0.detach().to(torch.float64)
        _set_initial(adj, fields)
        J = adj.forward(prob.nsteps)
        ctx.adj = adj
        ctx.dev = tt0.device
        ctx.pdev = _devices(re, theta)
        return torch.tensor(J, dtype=torch.float64, device=tt0.device)

    @staticmethod
    def backward(ctx, gJ):
        # the context belongs to prob (prob.adjoint caches one, reused while
        # Re and theta stay the same; prob.close() frees it)
        adj, ctx.adj = ctx.adj, None
        grads, dre, dth = _backward(adj, ["tt"], ctx.dev)
        g = grads["tt"][1:-1, 1:-1, 1:-1]
        dv_re, dv_th = ctx.pdev
        return (gJ * g, gJ.to(dv_re or ctx.dev) * _scalar(dre, dv_re or ctx.dev),
                gJ.to(dv_th or ctx.dev) * _scalar(dth, dv_th or ctx.dev), None)


class _StaticState(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, prob):
        arrs, tb = prob.static_numpy(float(theta))
        ctx.prob, ctx.theta, ctx.dev = prob, float(theta), theta.device
        dev = theta.device
        return tuple(torch.from_numpy(a).to(dev) for a in arrs) + (
            torch.tensor(tb, dtype=torch.float64, device=dev),)

    @staticmethod
    def backward(ctx, *g):
        th = ctx.theta
        h = 1e-6 * abs(th) if th != 0.0 else 1e-6
        ap, tbp = ctx.prob.static_numpy(th + h)
        am, tbm = ctx.prob.static_numpy(th - h)
        total = 0.0
        for gi, p_, m_ in zip(g[:5], ap, am):
            if gi is not None:
                dq = torch.from_numpy((p_ - m_) / (2 * h)).to(gi.device)
                total += float(torch.sum(gi * dq))
        if g[5] is not None:
            total += float(g[5]) * (tbp - tbm) / (2 * h)
        return torch.tensor(total, dtype=torch.float64, device=ctx.dev), None


class _KineticEnergyState(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ro0, tt0, bx0, by0, bz0, re, theta, tb, prob):
        cfg = cfg_with(prob.cfg, re=float(re), theta=float(theta), tb=float(tb))
        adj = Adjoint(cfg, prob.metrics, prob.every)
        adj.set_fixed_dt(prob.fixed_dt)
        fields = prob.base_on(tt0.device)
        sl = (slice(1, -1),) * 3
        ins = {"ro": ro0, "tt": tt0, "bx": bx0, "by": by0, "bz": bz0}
        for n in adj.names:
            if n in ins:
                fields[n][sl] = ins[n].detach().to(tt0.device, torch.float64)
        _set_initial(adj, fields)
        J = adj.forward(prob.nsteps)
        ctx.adj, ctx.dev = adj, tt0.device
        ctx.pdev = _devices(re, theta, tb)
        return torch.tensor(J, dtype=torch.float64, device=tt0.device)

    @staticmethod
    def backward(ctx, gJ):
        adj, ctx.adj = ctx.adj, None
        grads, dre, dth = _backward(adj, adj.names, ctx.dev)
        dtb = adj.dJ_dtb()
        adj.close()
        sl = (slice(1, -1),) * 3

        def g(n):
            if n not in grads:
                return None
            return gJ * grads[n][sl]

        def sc(v, d):
            d = d or ctx.dev
            return gJ.to(d) * _scalar(v, d)

        dv = ctx.pdev
        return (g("ro"), g("tt"), g("bx"), g("by"), g("bz"), sc(dre, dv[0]), sc(dth, dv[1]),
                sc(dtb, dv[2]), None)
