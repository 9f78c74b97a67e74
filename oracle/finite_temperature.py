"""Finite-temperature (iTPO) pipeline restated in numpy: the TPO initial
state, the simple update of the density operator and its single-layer CTM
measurements (TeNeS src/iTPS/tensors.cpp:140-203, simple_update.cpp:141-178,
finite_temperature.cpp:23-84, measure.cpp:110-143).
Test infrastructure: the checker of the GPU library, never part of it.

TPO site tensors (left, up, right, down, phys_ket, phys_bra).  A two-site
evolution operator U[in1, in2, out1, out2] acts on the ket legs and its
complex conjugate on the bra legs (mptensor::kron(U, conj(U)) on the merged
physical leg, phys_ket the slower index); for the real operators of the
test inputs the order of the merge does not matter.
"""
from __future__ import annotations

import numpy as np

from . import contract_density as KD
from . import ctm_single as CS
from . import itps
from . import transfer as TM
from .simple_update import Simple_update_bond


def initial_tensors_density(lat, pdims, vdims, dtype=float):
    """tensors.cpp:154-203: the identity operator on virtual index 0."""
    Tn = []
    for i in range(lat.N_UNIT):
        d, v = pdims[i], vdims[i]
        T = np.zeros((v[0], v[1], v[2], v[3], d, d), dtype=dtype)
        for p in range(d):
            T[0, 0, 0, 0, p, p] = 1.0
        Tn.append(T)
    return Tn


def _kron4(U, V):
    """mptensor::kron of two rank-4 tensors: leg k of the result is the
    merged (leg k of U, leg k of V), U's index the slower."""
    K = np.einsum("abcd,efgh->aebfcgdh", U, V)
    s = [U.shape[k] * V.shape[k] for k in range(4)]
    return K.reshape(s)


def simple_update_density(Tn, lambdas, lat, source, leg, op, inverse_lambda_cut=1e-12):
    """simple_update.cpp:141-178 for one evolution operator (one-site
    operators act as T <- (T . op) . conj(op) on the physical legs)."""
    if op.ndim == 2:
        T = np.tensordot(Tn[source], op, ([4], [0]))
        Tn[source] = np.tensordot(T, np.conj(op), ([4], [1]))
        return
    target = lat.neighbor(source, leg)
    tleg = (leg + 2) % 4
    a, b = Tn[source], Tn[target]
    sa, sb = a.shape, b.shape
    t1, t2, lam = Simple_update_bond(a.reshape(sa[:4] + (sa[4] * sa[5],)), b.reshape(sb[:4] + (sb[4] * sb[5],)),
                                     lambdas[source], lambdas[target], _kron4(op, np.conj(op)), leg,
                                     inverse_lambda_cut)
    lambdas[source][leg] = lam
    lambdas[target][tleg] = lam
    Tn[source] = t1.reshape(t1.shape[:4] + (op.shape[2], op.shape[2]))
    Tn[target] = t2.reshape(t2.shape[:4] + (op.shape[3], op.shape[3]))


def measure_density(Tn, lat, p, onesite, twosite, site_ops, r_max, pairs, beta, tm_params=None):
    """measure.cpp:110-143: a fresh single-layer CTM environment, then the
    one-site and two-site observables, the correlation functions and (for
    beta > 0) the correlation lengths.  Returns a dict of raw results."""
    env, count, _ = CS.Calc_CTM_Environment_density(Tn, p, lat)
    one, norms = itps.measure_onesite(env, Tn, lat, onesite, K=KD)
    two, _ = itps.measure_twosite(env, Tn, lat, twosite, onesite, K=KD)
    cor = itps.measure_correlation(env, Tn, lat, site_ops, r_max, pairs, K=KD) if r_max > 0 else {}
    xi = []
    if beta > 0.0:
        xi = itps.correlation_lengths(env, Tn, lat, tm_params or TM.TMParameters())
    return {"beta": beta, "env": env, "ctm_count": count, "onesite": one, "norms": norms,
            "twosite": two, "correlation": cor, "correlation_length": xi}


def finite_temperature(Tn, lat, evolution_ops, nsteps, tau, measure_interval, measure, inverse_lambda_cut=1e-12):
    """finite_temperature.cpp:23-84 for one simple-update group: measure at
    beta = 0, after every measure_interval steps and after the last step;
    beta grows by 2 tau per step.  measure(Tn, beta) returns one record."""
    vd = [[Tn[i].shape[l] for l in range(4)] for i in range(lat.N_UNIT)]
    lambdas = [[np.ones(vd[i][l]) for l in range(4)] for i in range(lat.N_UNIT)]
    beta = 0.0
    records = [measure(Tn, beta)]
    for step in range(nsteps):
        for source, leg, op in evolution_ops:
            simple_update_density(Tn, lambdas, lat, source, leg, op, inverse_lambda_cut)
        beta += tau * 2.0
        if (step + 1) % measure_interval == 0:
            records.append(measure(Tn, beta))
    if measure_interval == 0 or nsteps % measure_interval != 0:
        records.append(measure(Tn, beta))
    return records
