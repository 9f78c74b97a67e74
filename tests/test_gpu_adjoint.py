"""GPU bond energy and its adjoint (tenes_env_bond_energy) against the oracle
(oracle/adjoint.py, pinned in test_oracle_adjoint.py), central finite
differences and the torch.autograd harness (tenes_amd/adjoint.py)."""
from __future__ import annotations

import numpy as np
import pytest

import tenes_amd as ta
from ctm_cases import random_tensors
from oracle import adjoint as A
from oracle import ctm, models
from oracle.lattice import SquareLattice

pytestmark = pytest.mark.gpu

LAT = SquareLattice(2, 2)
KINDS = ("C1", "C2", "C3", "C4", "eTt", "eTr", "eTb", "eTl")


def _bonds():
    H = models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    return [(s, 1, 0, H) for s in range(4)] + [(s, 0, 1, H) for s in range(4)]


def _download_env(genv):
    env = ctm.Env(4)
    for k in KINDS:
        for i in range(4):
            getattr(env, k)[i] = genv.get(k, i).numpy()
    return env


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_bond_energy_and_gradient_vs_oracle_small(gpu_ctx):
    """D=3, chi=12 oracle environment uploaded to the GPU; every bond
    orientation (sources on both sides), energy and gradient vs the oracle."""
    Tn = random_tensors(3, 2, seed=21)
    p = ctm.PEPSParameters(CHI=12, Max_CTM_Iteration=3, CTM_Convergence_Epsilon=0.0)
    env, _, _ = ctm.Calc_CTM_Environment(list(Tn), p, LAT)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    for k in KINDS:
        for i in range(4):
            genv.set(k, i, getattr(env, k)[i])
    H = _bonds()[0][3]
    rng = np.random.default_rng(2)
    op = rng.standard_normal((2, 2, 2, 2))
    bonds = [(0, 1, 0, H), (1, -1, 0, op), (2, 0, 1, op), (3, 0, -1, H), (1, 1, 0, op), (0, 0, -1, op)]
    E_ref, g_ref = A.energy_and_grad(env, list(Tn), LAT, bonds)
    E, g = genv.bond_energy(bonds, grad=True)
    assert abs(E - E_ref) < 1e-12 * max(1.0, abs(E_ref)), (E, E_ref)
    for i in range(4):
        assert _rel(g[i], g_ref[i]) < 1e-11, (i, _rel(g[i], g_ref[i]))
    E2, none = genv.bond_energy(bonds)
    assert none is None and E2 == E


def test_heisenberg_d4_chi64_gradient(gpu_ctx):
    """Heisenberg, 2x2, D=4, chi=64, unit-norm random tensors, GPU CTM
    environment (20 sweeps) held fixed: dE/dTn for all four site tensors
    against the CPU FP64 oracle on the same environment (1e-10) and against
    central differences of the GPU energy along 32 random directions
    (h = 1e-6, agreement to 1e-6 relative)."""
    from tenes_amd import inputs

    Tn = inputs.random_tensors(4, 4, 2, seed=11)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    genv.calc_ctm_environment(ta.peps_parameters(CHI=64, Max_CTM_Iteration=20, CTM_Convergence_Epsilon=0.0))
    bonds = _bonds()
    E, g = genv.bond_energy(bonds, grad=True)
    env = _download_env(genv)
    E_ref, g_ref = A.energy_and_grad(env, list(Tn), LAT, bonds)
    assert abs(E - E_ref) < 1e-12 * abs(E_ref), (E, E_ref)
    for i in range(4):
        assert _rel(g[i], g_ref[i]) < 1e-10, (i, _rel(g[i], g_ref[i]))
    rng = np.random.default_rng(3)
    h = 1e-6
    worst = 0.0
    for _ in range(32):
        V = [rng.standard_normal(t.shape) for t in Tn]
        genv.set_site_tensors([t + h * v for t, v in zip(Tn, V)])
        Ep, _ = genv.bond_energy(bonds)
        genv.set_site_tensors([t - h * v for t, v in zip(Tn, V)])
        Em, _ = genv.bond_energy(bonds)
        fd = (Ep - Em) / (2 * h)
        an = sum(float(np.sum(x * v)) for x, v in zip(g, V))
        worst = max(worst, abs(fd - an) / abs(an))
    genv.set_site_tensors(Tn)
    print(f"E = {E:.15f}; worst FD rel diff over 32 directions {worst:.2e}")
    assert worst < 1e-6, worst


def test_heisenberg_d4_chi64_gradient_fp32_contractions(gpu_ctx):
    """The FP32 variant: the same energy and gradient with the context's
    contractions on the FP32 MFMA (tenes_ctx_set_fp32_contractions), the
    environment held fixed.  Error against the FP64 gradient: 1e-5 relative
    per site tensor (measured 1.4e-6); energy 1e-6 absolute (measured 1.9e-8:
    the random tensors' energy per site, -1.7e-3, is a cancellation of bond
    terms of order 0.1, so a relative measure would grade the cancellation,
    not the contraction)."""
    from tenes_amd import inputs

    Tn = inputs.random_tensors(4, 4, 2, seed=11)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    genv.calc_ctm_environment(ta.peps_parameters(CHI=64, Max_CTM_Iteration=20, CTM_Convergence_Epsilon=0.0))
    bonds = _bonds()
    E, g = genv.bond_energy(bonds, grad=True)
    gpu_ctx.set_fp32_contractions(True)
    try:
        E32, g32 = genv.bond_energy(bonds, grad=True)
    finally:
        gpu_ctx.set_fp32_contractions(False)
    errs = [_rel(g32[i], g[i]) for i in range(4)]
    print(f"FP32 contractions: energy abs err {abs(E32 - E):.2e}, gradient rel err {max(errs):.2e}")
    assert abs(E32 - E) < 1e-6
    assert max(errs) < 1e-5, errs
    assert max(errs) > 0.0  # the FP32 path really ran


def test_autograd_and_gradient_descent(gpu_ctx):
    """torch.autograd.Function: .grad equals the C-ABI gradient; then 50
    normalised gradient-descent steps (lr 0.1, environment recomputed every
    step from the previous one) lower the Heisenberg energy of D=4, chi=64
    random tensors from ~0 to below -0.5 (measured: -0.0017 -> -0.591 at
    step 20 -> -0.576 at step 50; the fixed-environment gradient leaves out
    the environment's own dependence on the tensors, so the descent is not
    monotone once the environment moves)."""
    import torch

    from tenes_amd import adjoint, inputs

    Tn = inputs.random_tensors(4, 4, 2, seed=11)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    prm = ta.peps_parameters(CHI=64, Max_CTM_Iteration=10, CTM_Convergence_Epsilon=1e-9)
    genv.calc_ctm_environment(prm)
    bonds = _bonds()
    E_ref, g_ref = genv.bond_energy(bonds, grad=True)
    T = [torch.tensor(t, requires_grad=True) for t in Tn]
    E = adjoint.BondEnergy.apply(genv, bonds, *T)
    E.backward()
    assert abs(E.item() - E_ref) < 1e-14 * abs(E_ref)
    for t, gr in zip(T, g_ref):
        assert _rel(t.grad.numpy(), gr) < 1e-14
    energies, _ = adjoint.gradient_descent(genv, Tn, bonds, steps=50, lr=0.1, ctm_params=prm)
    print("energies: " + " ".join(f"{e:.6f}" for e in energies))
    assert len(energies) == 50
    assert energies[-1] < -0.5 and energies[0] > -0.1, energies
    assert all(b < a for a, b in zip(energies[:15], energies[1:15])), energies
