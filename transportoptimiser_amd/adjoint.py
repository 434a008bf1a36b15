"""Adjoint of the MTE simulation: gradient of the mean total end population
with respect to crossing mortality and growth rates (roadmc_mte_adjoint_*),
and config 4's objective with the smoothed species-state value added
(roadmc_rovs_smooth_grad)."""
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


def rov_species_on(sd, flow, params):
    """The species-state valuation's species on an MTE species' landscape
    (rov_species.species_on)."""
    from .rov_species import species_on
    return species_on(sd, flow, params)


def objective_gradient(ctx, problem: RoadProblem, rov_problem, policy, tau: float,
                       rov_weight: float = 1.0, checkpoint_every: int = 10):
    """Config 4's objective with a non-zero ROV term:
    J = mean total end population of ``problem`` (one species) + rov_weight x
    the smoothed-validity species-state value of ``rov_problem`` under
    ``policy`` [T][N] (validity prod_s sigmoid((A_s - threshold_s) / tau)),
    whose species 0 is ``rov_species_on(problem.species[0], ...)``.  Returns
    (J, dJ/dm (P, P), dJ/dr (P,)) with m the crossing mortality of the MTE
    species (surv = 1 - m, rows the origin): the MTE adjoint plus rov_weight x
    dV/dm = -sum_c (q_c / q_max) dV/dsurv_c."""
    from .rov_species import smooth_value_gradient
    obj, gm, gr = mte_gradient(ctx, problem, checkpoint_every)
    v, gs = smooth_value_gradient(ctx, rov_problem, policy, tau)
    q = np.asarray(rov_problem.flow, dtype=np.float64)
    q = q / q.max()
    g_rov = -np.tensordot(q, gs[0], axes=(0, 0))
    return obj + rov_weight * v, gm + rov_weight * g_rov, gr

