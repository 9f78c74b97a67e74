"""Finite-temperature (iTPO) driver over the GPU library: TeNeS's
iTPS::finite_temperature (src/iTPS/finite_temperature.cpp:23-84) with the
TPO initial state of initialize_tensors_density (tensors.cpp:140-203) and
the measurements of measure_density (measure.cpp:110-143): the single-layer
CTM environment (tenes_calc_ctm_environment_density), one-site and two-site
observables, the energy density, correlation functions and, for beta > 0,
correlation lengths.

Operators use the dict formats of the measurement drivers:
  onesite: {"group", "site", "op" (d, d), "coeff"}
  twosite: {"group", "source", "dx", "dy", "coeff",
            "op" (d1, d2, d1, d2) or "ops" [g_source, g_target]}
Complex operators are evaluated by linearity from their real and imaginary
parts (the site tensors are real).
"""
from __future__ import annotations

from math import gcd

import numpy as np


def initial_tpo(pdims, vdims):
    """tensors.cpp:154-203: the identity operator on virtual index 0."""
    out = []
    for d, v in zip(pdims, vdims):
        T = np.zeros((v[0], v[1], v[2], v[3], d, d))
        for p in range(d):
            T[0, 0, 0, 0, p, p] = 1.0
        out.append(T)
    return out


def _parts(op):
    """[(real part, coefficient)] of a possibly complex operator."""
    op = np.asarray(op)
    parts = [(np.ascontiguousarray(op.real, dtype=np.float64), 1.0)]
    if np.iscomplexobj(op) and np.any(op.imag != 0):
        parts.append((np.ascontiguousarray(op.imag, dtype=np.float64), 1j))
    return parts


def _noskew(LX, LY, skew):
    skew %= LX
    if skew == 0:
        return LX, LY
    return LX, LY * ((LX // gcd(LX, skew) * skew) // skew)


def measure(env, LX, LY, skew, onesite, twosite, site_ops, r_max, pairs, params, beta, tm_params=None):
    """measure.cpp:110-143 on the GPU.  Returns a record with the keys of
    the oracle's measure_density: onesite {group: {site: value}}, twosite
    {group: {(source, dx, dy): value}}, correlation {(left_group, left,
    right_group, dx, dy): value}, correlation_length [(dir, coord, xi,
    [e_i], eigenvalues)], ctm_count."""
    count, _ = env.calc_ctm_environment_density(params)
    N = LX * LY
    norms = [env.contract_one_site(i) for i in range(N)]
    one = {}
    for o in onesite:
        i = o["site"]
        v = sum(c * env.contract_one_site(i, p) for p, c in _parts(o["op"]))
        one.setdefault(o["group"], {})[i] = o.get("coeff", 1.0) * v / norms[i]
    two = {}
    for o in twosite:
        s, dx, dy = o["source"], o["dx"], o["dy"]
        norm = env.contract_cluster(s, dx, dy)
        if "op" in o:
            v = sum(c * env.contract_two_sites_op12(s, dx, dy, p) for p, c in _parts(o["op"]))
        else:
            t = _other(LX, LY, skew, s, dx, dy)
            A, B = site_ops[(s, o["ops"][0])], site_ops[(t, o["ops"][1])]
            v = sum(ca * cb * env.contract_cluster(s, dx, dy, pa, pb)
                    for pa, ca in _parts(A) for pb, cb in _parts(B))
        two.setdefault(o["group"], {})[(s, dx, dy)] = o.get("coeff", 1.0) * v / norm
    cor = {}
    if r_max > 0:
        rg_of = {}
        for lg, rg in pairs:
            rg_of.setdefault(lg, []).append(rg)
        for left in range(N):
            for lg, rgs in rg_of.items():
                if (left, lg) not in site_ops:
                    continue
                for direction in (0, 1):
                    # right operators are taken from the left site's groups,
                    # as the one-site operator set is the same on every site
                    for rg in rgs:
                        L, R = site_ops[(left, lg)], site_ops[(left, rg)]
                        val = sum(cl * cr * env.correlation(left, pl, [pr], r_max, direction)[:, 0]
                                  for pl, cl in _parts(L) for pr, cr in _parts(R))
                        for r in range(r_max):
                            key = (lg, left, rg, r + 1, 0) if direction == 0 else (lg, left, rg, 0, r + 1)
                            cor[key] = val[r]
    xi = []
    if beta > 0.0:
        LXn, LYn = _noskew(LX, LY, skew)
        for d in (0, 1):
            W = LY if d == 0 else LX
            L = LXn if d == 0 else LYn
            for fixed in range(W):
                ev = env.transfer_matrix_eigenvalues(d, fixed, tm_params)
                if len(ev) == 1:
                    xi.append((d, fixed, 0.0, [], ev))
                    continue
                e0 = abs(ev[0])
                xi.append((d, fixed, -L / np.log(abs(ev[1]) / e0),
                           [-np.log(abs(v) / e0) / L for v in ev[1:]], ev))
    return {"beta": beta, "ctm_count": count, "onesite": one, "norms": norms, "twosite": two,
            "correlation": cor, "correlation_length": xi}


def _other(LX, LY, skew, i, dx, dy):
    """SquareLattice::other (src/SquareLattice.hpp:54-163): the site (dx, dy)
    away from i; crossing the top boundary upwards shifts x by -skew
    (T(x, y) = T(x + skew, y + LY))."""
    x, y = i % LX + dx, i // LX + dy
    k = y // LY
    return (x - k * skew) % LX + LX * (y - k * LY)


def density(onesite_vals, twosite_vals, pdims, onesite_names, twosite_names):
    """density.cpp:39-76: per-site averages and the energy density."""
    nsites = sum(1 for d in pdims if d > 1)
    res, energy = {}, 0.0
    for g, name in onesite_names.items():
        v = sum(val for i, val in onesite_vals.get(g, {}).items() if pdims[i] > 1)
        res[name] = v / nsites
        if name == "site_hamiltonian":
            energy += np.real(v)
    for g, name in twosite_names.items():
        v = sum(twosite_vals.get(g, {}).values())
        res[name] = v / nsites
        if name == "bond_hamiltonian":
            energy += np.real(v)
    res["Energy"] = energy / nsites
    return res


def run(env, pdims, vdims, evolution, nsteps, tau, interval, measure_fn, inverse_lambda_cut=1e-12):
    """finite_temperature.cpp:23-84 for one simple-update group: measure at
    beta = 0, after every `interval` steps and after the last one; beta
    grows by 2 tau per step.  evolution = [(source, leg, U (d1, d2, d1,
    d2))]; measure_fn(env, beta) returns one record."""
    env.set_site_tensors(initial_tpo(pdims, vdims))
    beta = 0.0
    records = [measure_fn(env, beta)]
    done = 0
    while done < nsteps:
        n = min(interval, nsteps - done) if interval > 0 else nsteps - done
        env.simple_update(evolution, n, inverse_lambda_cut)
        done += n
        beta += 2.0 * tau * n
        if interval > 0 and done % interval == 0:
            records.append(measure_fn(env, beta))
    if interval == 0 or nsteps % interval != 0:
        records.append(measure_fn(env, beta))
    return records
