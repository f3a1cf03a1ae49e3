"""torch.autograd binding of the discrete adjoint (include/r3d_adj.h).

KineticEnergyRun(args, nsteps) is the differentiable map
    (tt0, Re, theta) -> J = final kinetic energy of a hydrodynamic run,
where tt0 is the interior initial temperature (npz, npy, npx) replacing that
of the run's initial state (built by the host setup from the reference CLI
flags `args`; hydrodynamic or MHD).  The forward is the plain bit-exact time step on the GPU; the
backward is the hand-written adjoint.  Composing tt0 = T_static * (1 + delta)
in torch gives the gradient with respect to the perturbation field delta.
"""
import numpy as np
import torch

from . import Adjoint, PRIMARIES, Run, cfg_with


class KineticEnergyRun:
    def __init__(self, args, nsteps, checkpoint_every=20):
        self.run = Run(args)
        self.base = dict(self.run.state())
        self.metrics = self.run.metrics()
        self.cfg = self.run.cfg()
        self.nsteps = nsteps
        self.every = checkpoint_every
        self._key = None
        self._adj = None

    def adjoint(self, re, theta):
        key = (re, theta)
        if key != self._key:
            if self._adj is not None:
                self._adj.close()
            self._adj = Adjoint(cfg_with(self.cfg, re=re, theta=theta), self.metrics, self.every)
            self._key = key
        return self._adj

    def initial_tt(self):
        """The run's own interior initial temperature."""
        return torch.from_numpy(self.base["tt"][1:-1, 1:-1, 1:-1].copy())

    def __call__(self, tt0, re, theta):
        return _KineticEnergy.apply(tt0, re, theta, self)


class _KineticEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tt0, re, theta, prob):
        adj = prob.adjoint(float(re), float(theta))
        fields = {k: v.copy() for k, v in prob.base.items()}
        fields["tt"][1:-1, 1:-1, 1:-1] = tt0.detach().to("cpu", torch.float64).numpy()
        adj.set_initial(fields)
        J = adj.forward(prob.nsteps)
        ctx.adj = adj
        ctx.dev = tt0.device
        return torch.tensor(J, dtype=torch.float64, device=tt0.device)

    @staticmethod
    def backward(ctx, gJ):
        grads, dre, dth = ctx.adj.backward()
        g = torch.from_numpy(grads["tt"][1:-1, 1:-1, 1:-1].copy()).to(ctx.dev)
        return (gJ * g, gJ * torch.tensor(dre, dtype=torch.float64, device=ctx.dev),
                gJ * torch.tensor(dth, dtype=torch.float64, device=ctx.dev), None)
