"""Adjoint of the MTE simulation: gradient of the mean total end population
with respect to crossing mortality and growth rates (roadmc_mte_adjoint_*)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as nat
from .api import _Marshal
from .landscape import RoadProblem


class Adjoint:
    """Forward pass with checkpoints; backward() returns the gradients."""

    def __init__(self, ctx, problem: RoadProblem, checkpoint_every: int = 10):
        self.ctx = ctx
        self.P = problem.species[0].n_patches
        m = _Marshal([problem])
        h = C.c_void_p()
        obj = C.c_double()
        nat.check(nat.lib().roadmc_mte_adjoint_fwd(ctx.handle, m.problems, checkpoint_every,
                                                   C.byref(h), C.byref(obj)), ctx.handle)
        self.handle = h
        self.objective = obj.value

    def backward(self, d_objective: float = 1.0):
        gm = np.zeros(self.P * self.P)
        gr = np.zeros(self.P)
        dp = C.POINTER(C.c_double)
        nat.check(nat.lib().roadmc_mte_adjoint_bwd(self.handle, float(d_objective),
                                                   gm.ctypes.data_as(dp), gr.ctypes.data_as(dp)),
                  self.ctx.handle)
        return gm.reshape(self.P, self.P, order="F"), gr

    def close(self):
        if self.handle:
            nat.lib().roadmc_adjoint_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def mte_gradient(ctx, problem: RoadProblem, checkpoint_every: int = 10):
    """(objective, dJ/dm (P, P), dJ/dr (P,)) for one species."""
    with Adjoint(ctx, problem, checkpoint_every) as a:
        gm, gr = a.backward(1.0)
        return a.objective, gm, gr
