"""k_eff of a 2D deck as a differentiable function of the per-band
relaxation-time scale factors s_b (tau_b -> s_b tau_b, lattice ratios
following) and of the hot wall's temperature: a torch.autograd.Function
whose forward runs the GPU solve to the given tolerances and whose backward
returns the discrete-adjoint gradient (gbte_adjoint).

The adjoint works at the converged fixed point, so it needs no stored
iterates: the forward keeps nothing but the two gradient vectors."""
from __future__ import annotations

import dataclasses
import os
import tempfile

import numpy as np

from . import BTESolver, Problem, decks


def keff_and_grad(spec: decks.DeckSpec, tau_scale, t_hot: float, hot_index: int, axis: int = 0,
                  delta_t: float = 2.0, max_iter: int = 100000, tol_temp: float = 1e-15, tol_flux: float = 1e-13,
                  adjoint_tol: float = 1e-13, workdir: str | None = None):
    """(k_eff, dk/ds [n_bands], dk/dT_hot, solve stats, adjoint stats) of
    `spec` with tau scaled by tau_scale and boundary condition hot_index (an
    isothermal wall) at t_hot.  dk/dT_hot sums the adjoint over the faces of
    that wall (the type-1 faces at temperature t_hot)."""
    s = np.asarray(tau_scale, dtype=np.float64)
    sp = dataclasses.replace(spec, bands=decks.scale_tau(spec.bands, s))
    sp = decks.set_wall_temperature(sp, hot_index, float(t_hot))
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        p = Problem(decks.write_deck(os.path.join(td, "deck"), sp))
        sol = BTESolver(p)
        try:
            sol.fourier()
            st = sol.solve(max_iter, tol_temp, tol_flux)
            k = sol.keff(axis, delta_t)
            g, gtb, ast = sol.adjoint(axis, delta_t, max_iter, adjoint_tol)
            bt, btemp = p.array("bound_type"), p.array("bound_temp")
            g_hot = float(gtb[(bt == 1) & (btemp == float(t_hot))].sum())
        finally:
            sol.close()
            p.close()
    # the adjoint's derivative is along the deck's own tau: d/ds_b = (d/ds'_b) / s_b
    return k, g / s, g_hot, st, ast


try:
    import torch

    class KeffFunction(torch.autograd.Function):
        """k = KeffFunction.apply(tau_scale, t_hot, spec, hot_index, axis, delta_t):
        tau_scale a 1-D tensor of n_bands factors, t_hot a 0-d tensor (K)."""

        @staticmethod
        def forward(ctx, tau_scale, t_hot, spec, hot_index, axis=0, delta_t=2.0):
            k, g, g_hot, _, _ = keff_and_grad(spec, tau_scale.detach().cpu().double().numpy(),
                                              float(t_hot.detach().cpu()), hot_index, axis, delta_t)
            ctx.save_for_backward(torch.as_tensor(g, dtype=tau_scale.dtype),
                                  torch.as_tensor(g_hot, dtype=t_hot.dtype))
            return torch.as_tensor(k, dtype=tau_scale.dtype)

        @staticmethod
        def backward(ctx, grad_k):
            g, g_hot = ctx.saved_tensors
            return grad_k * g, grad_k * g_hot, None, None, None, None

    def keff(tau_scale, t_hot, spec, hot_index, axis: int = 0, delta_t: float = 2.0):
        """Differentiable k_eff (see KeffFunction)."""
        return KeffFunction.apply(tau_scale, t_hot, spec, hot_index, axis, delta_t)

except ImportError:  # torch absent: keff_and_grad still works
    pass
