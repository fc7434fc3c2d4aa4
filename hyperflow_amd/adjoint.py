"""Sensitivities of fixed-step chemistry (no reference counterpart).

The integration is the V3 reaction step's right-hand side
(HyperFlow_V3/src/chemistry.cpp:36-53) advanced by ``nsteps`` backward-Euler
steps of size ``h`` at fixed density and internal energy e = e(T0, Y0). The
objective per cell is

    J = T(t_end) + sum_n h qdot(y_{n+1}),   qdot = -sum_k u_k wdot_k  [W/m^3]

``ChemSensitivity.adjoint`` returns dJ/dlnA_i (A_i multiplies reaction i's
forward and reverse rates, k0 and k_inf alike), dJ/dT0 and dJ/dY0 for every
cell from one reverse sweep over checkpointed segments; ``tangent`` returns
the directional derivative along given directions (forward mode).
``FixedStepChemistry`` wraps the pair as a ``torch.autograd.Function`` of a
shared ln A shift and the per-cell initial T and Y.
"""
from __future__ import annotations

from ctypes import c_double

import numpy as np

from ._lib import check, load


def _addr(x):
    if x is None:
        return None
    try:
        import torch
        if isinstance(x, torch.Tensor):
            if x.dtype != torch.float64 or not x.is_contiguous():
                raise TypeError("expected a contiguous float64 tensor")
            return x.data_ptr()
    except ImportError:
        pass
    if isinstance(x, np.ndarray):
        if x.dtype != np.float64 or not x.flags.c_contiguous:
            raise TypeError("expected a contiguous float64 array")
        return x.ctypes.data
    raise TypeError(f"unsupported array type {type(x)}")


def _like(x, shape):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return torch.empty(shape, dtype=torch.float64, device=x.device)
    except ImportError:
        pass
    return np.empty(shape, dtype=np.float64)


class ChemSensitivity:
    """Adjoint and tangent of fixed-step chemistry for one mechanism.

    Arrays: rho (n,), T0 (n,), Y0 (n, ns) C-contiguous; numpy (host) or torch
    (device) float64, all of one kind. lnA_shift: optional (nr,) shift of
    every cell's ln A.
    """

    def __init__(self, mechanism):
        self.mech = mechanism
        self.ns, self.nr = mechanism.ns, mechanism.nr

    def adjoint(self, rho, T0, Y0, h, nsteps, ckpt_every=20, lnA_shift=None):
        n = int(rho.shape[0])
        J, g_lnA, g_T0, g_Y0 = _like(rho, (n,)), _like(rho, (n, self.nr)), _like(rho, (n,)), _like(rho, (n, self.ns))
        check(self.mech._h, load().hf_chem_adjoint(self.mech._h, n, _addr(rho), _addr(T0), _addr(Y0), c_double(h),
                                                   int(nsteps), int(ckpt_every), _addr(lnA_shift), _addr(J),
                                                   _addr(g_lnA), _addr(g_T0), _addr(g_Y0)))
        return J, g_lnA, g_T0, g_Y0

    def tangent(self, rho, T0, Y0, h, nsteps, v_lnA, v_T0, v_Y0, lnA_shift=None):
        n = int(rho.shape[0])
        J, dJ = _like(rho, (n,)), _like(rho, (n,))
        check(self.mech._h, load().hf_chem_tangent(self.mech._h, n, _addr(rho), _addr(T0), _addr(Y0), c_double(h),
                                                   int(nsteps), _addr(lnA_shift), _addr(v_lnA), _addr(v_T0),
                                                   _addr(v_Y0), _addr(J), _addr(dJ)))
        return J, dJ


def _function():
    import torch

    class FixedStepChemistry(torch.autograd.Function):
        """J(lnA_shift, T0, Y0) per cell, differentiable through the discrete adjoint.

        forward(sens, lnA_shift (nr,), T0 (n,), Y0 (n, ns), rho (n,), h, nsteps, ckpt_every) -> J (n,)
        """

        @staticmethod
        def forward(ctx, sens, lnA_shift, T0, Y0, rho, h, nsteps, ckpt_every):
            lnA_shift, T0, Y0, rho = (x.detach().contiguous() for x in (lnA_shift, T0, Y0, rho))
            J, g_lnA, g_T0, g_Y0 = sens.adjoint(rho, T0, Y0, h, nsteps, ckpt_every, lnA_shift)
            ctx.save_for_backward(g_lnA, g_T0, g_Y0)
            return J

        @staticmethod
        def backward(ctx, gJ):
            g_lnA, g_T0, g_Y0 = ctx.saved_tensors
            return (None, gJ @ g_lnA, gJ * g_T0, gJ[:, None] * g_Y0, None, None, None, None)

    return FixedStepChemistry


_FN = None


def fixed_step_chemistry(sens, lnA_shift, T0, Y0, rho, h, nsteps, ckpt_every=20):
    """Differentiable J per cell (torch tensors on the device of the mechanism)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(sens, lnA_shift, T0, Y0, rho, h, nsteps, ckpt_every)
