"""dE/dTn behind torch.autograd, over the C ABI (tenes_env_bond_energy).

`BondEnergy.apply(env, bonds, *Tn)` returns the energy per site of the
nearest-neighbour bond operators `bonds` = [(source, dx, dy, op12)] for the
site tensors Tn (float64 torch tensors, one per unit-cell site) with the CTM
environment held in `env` fixed.  The forward call computes the energy and
the gradient together on the GPU (the adjoint of every bond's two-site
reduced density matrix, recomputed bond by bond instead of stored); backward
only scales the stored gradient.

The torch side holds host tensors: the site tensors (a few KB) are uploaded
per call and the gradient comes back from the library.  torch's wheel
carries its own HIP runtime, and two HIP runtimes cannot both open the GPU in
one process, so the GPU belongs to libtenes_amd here and torch does only the
autograd bookkeeping.  The reference has no gradient (TeNeS
optimises by imaginary-time evolution); the energy is its bond_hamiltonian
energy (density.cpp:39-76 over twosite_obs.cpp:146-262 bond values).

`gradient_descent` is the optimisation loop around it: per step, the CTM
environment is recomputed for the current tensors (continuing from the
previous one), then one normalised gradient step is taken per tensor.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .core import Environment


class BondEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, env: Environment, bonds, *Tn):
        if len(Tn) != env.N:
            raise ValueError(f"{env.N} site tensors expected, got {len(Tn)}")
        for t in Tn:
            if t.dtype != torch.float64 or t.device.type != "cpu":
                raise TypeError("site tensors must be float64 host tensors (the GPU work runs in libtenes_amd)")
        env.set_site_tensors([t.detach().numpy() for t in Tn])
        E, grads = env.bond_energy(bonds, grad=True)
        ctx.save_for_backward(*[torch.from_numpy(g) for g in grads])
        return torch.tensor(E, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        return (None, None) + tuple(gE * g for g in ctx.saved_tensors)


def gradient_descent(env: Environment, Tn: Sequence[np.ndarray], bonds, steps: int, lr: float,
                     ctm_params):
    """`steps` steps of T_i <- (T_i - lr * dE/dT_i) / |...| (the energy does
    not depend on the norm of a site tensor, so each stays at unit norm).
    Returns (energies before each step, final tensors as numpy arrays)."""
    T = [torch.tensor(np.asarray(t, dtype=np.float64)) for t in Tn]
    for t in T:
        t /= torch.linalg.vector_norm(t)
        t.requires_grad_(True)
    energies = []
    for it in range(steps):
        env.set_site_tensors([t.detach().numpy() for t in T])
        env.calc_ctm_environment(ctm_params, initialize=(it == 0))
        E = BondEnergy.apply(env, bonds, *T)
        E.backward()
        energies.append(float(E.item()))
        with torch.no_grad():
            for t in T:
                t -= lr * t.grad
                t /= torch.linalg.vector_norm(t)
                t.grad = None
    return energies, [t.detach().numpy() for t in T]
