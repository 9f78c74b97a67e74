"""Measurement drivers restated in numpy: one-site and two-site observables
and the energy density (TeNeS src/iTPS/onesite_obs.cpp:35-123,
twosite_obs.cpp:31-288, density.cpp:30-140) and the correlation length
(correlation_length.cpp:24-124) and the correlation functions
(correlation_function.cpp:44-220).  The contraction module K is
oracle.contract (iTPS, double layer) or oracle.contract_density (iTPO,
single layer; TPO site tensors with two physical legs).  Operators are given as plain dicts:

  onesite: {"name", "group", "site", "op" (d, d), "coeff"}
  twosite: {"name", "group", "source", "dx", "dy", "coeff",
            "op" (d1, d2, d1, d2) or "ops" (one-site groups [g_source, g_target])}
"""
from __future__ import annotations

import numpy as np

from . import contract as K
from . import transfer as TM


def measure_onesite(env, Tn, lat, onesite_ops, K=K):
    """onesite_obs.cpp:85-98.  Returns ({group: {site: value}}, norms)."""
    norms = []
    for i in range(lat.N_UNIT):
        d = Tn[i].shape[4]
        norms.append(K.Contract_one_site_iTPS_CTM(env.C1[i], env.C2[i], env.C3[i], env.C4[i], env.eTt[i],
                                                  env.eTr[i], env.eTb[i], env.eTl[i], Tn[i], np.eye(d)))
    out = {}
    for op in onesite_ops:
        i = op["site"]
        val = K.Contract_one_site_iTPS_CTM(env.C1[i], env.C2[i], env.C3[i], env.C4[i], env.eTt[i],
                                           env.eTr[i], env.eTb[i], env.eTl[i], Tn[i], op["op"])
        out.setdefault(op["group"], {})[i] = op.get("coeff", 1.0) * val / norms[i]
    return out, norms


def cluster(lat, source, dx, dy):
    """indices[row][col] of twosite_obs.cpp:119-176."""
    ncol = abs(dx) + 1
    nrow = abs(dy) + 1
    source_col, target_col = (0, ncol - 1) if dx >= 0 else (ncol - 1, 0)
    source_row, target_row = (nrow - 1, 0) if dy >= 0 else (0, nrow - 1)
    idx = [[lat.other(source, col - source_col, source_row - row) for col in range(ncol)]
           for row in range(nrow)]
    return nrow, ncol, source_row, source_col, target_row, target_col, idx


def _cluster_args(env, idx, nrow, ncol):
    C = [env.C1[idx[0][0]], env.C2[idx[0][ncol - 1]], env.C3[idx[nrow - 1][ncol - 1]], env.C4[idx[nrow - 1][0]]]
    eTt = [env.eTt[idx[0][c]] for c in range(ncol)]
    eTb = [env.eTb[idx[nrow - 1][c]] for c in range(ncol)]
    eTl = [env.eTl[idx[r][0]] for r in range(nrow)]
    eTr = [env.eTr[idx[r][ncol - 1]] for r in range(nrow)]
    return C, eTt, eTr, eTb, eTl


def measure_twosite(env, Tn, lat, twosite_ops, onesite_ops, K=K):
    """twosite_obs.cpp:46-262.  Returns ({group: {(source, dx, dy): value}}, norms)."""
    site_op = {(o["site"], o["group"]): o["op"] for o in onesite_ops}
    norms = {}
    out = {}
    for op in twosite_ops:
        source, dx, dy = op["source"], op["dx"], op["dy"]
        nrow, ncol, srow, scol, trow, tcol, idx = cluster(lat, source, dx, dy)
        if ncol > 4 or nrow > 4:
            continue
        Tn_ = [[Tn[idx[r][c]] for c in range(ncol)] for r in range(nrow)]
        ops_ = [[np.eye(Tn[idx[r][c]].shape[4]) for c in range(ncol)] for r in range(nrow)]
        C, eTt, eTr, eTb, eTl = _cluster_args(env, idx, nrow, ncol)
        key = (idx[nrow - 1][0], nrow - 1, ncol - 1)
        if key not in norms:
            norms[key] = K.Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, Tn_, ops_)
        norm = norms[key]
        if "op" in op and nrow * ncol != 2:
            # twosite_obs.cpp:223-243: op = sum_k s_k U_k (x) V_k from the SVD
            # over (source legs 0, 2 | target legs 1, 3), one cluster
            # contraction per product of one-site operators
            o = op["op"]
            d1, d2 = o.shape[0], o.shape[1]
            M = o.transpose(0, 2, 1, 3).reshape(d1 * d1, d2 * d2)
            U, s, VT = np.linalg.svd(M, full_matrices=False)
            value = 0.0
            for k in range(len(s)):
                ops_[srow][scol] = U[:, k].reshape(d1, d1)
                ops_[trow][tcol] = VT[k].reshape(d2, d2)
                value += s[k] * K.Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, Tn_, ops_)
        elif "op" in op:
            o = op["op"]
            if nrow == 2:
                top, bottom = idx[0][0], idx[1][0]
                o = o if top == source else o.transpose(1, 0, 3, 2)
                value = K.Contract_two_sites_vertical_op12_iTPS_CTM(
                    env.C1[top], env.C2[top], env.C3[bottom], env.C4[bottom], env.eTt[top], env.eTr[top],
                    env.eTr[bottom], env.eTb[bottom], env.eTl[bottom], env.eTl[top], Tn[top], Tn[bottom], o)
            else:
                left, right = idx[0][0], idx[0][1]
                o = o if left == source else o.transpose(1, 0, 3, 2)
                value = K.Contract_two_sites_horizontal_op12_iTPS_CTM(
                    env.C1[left], env.C2[right], env.C3[right], env.C4[left], env.eTt[left], env.eTt[right],
                    env.eTr[right], env.eTb[right], env.eTb[left], env.eTl[left], Tn[left], Tn[right], o)
        else:
            g0, g1 = op["ops"]
            ops_[srow][scol] = site_op[(source, g0)]
            target = lat.other(source, dx, dy)
            ops_[trow][tcol] = site_op[(target, g1)]
            value = K.Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, Tn_, ops_)
        out.setdefault(op["group"], {})[(source, dx, dy)] = op.get("coeff", 1.0) * value / norm
    return out, norms


def density(onesite_vals, twosite_vals, lat, pdims, onesite_names, twosite_names):
    """density.cpp:39-76: per-site averages and the energy
    (site_hamiltonian + bond_hamiltonian groups)."""
    nsites = sum(1 for i in range(lat.N_UNIT) if pdims[i] > 1)
    res = {}
    energy = 0.0
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


def measure_multisite(env, Tn, lat, site_ops, multisite_ops, K=K):
    """multisite_obs.cpp:46-220 (CTM environment): multisite_ops =
    [{"group", "source", "dx": [...], "dy": [...], "ops": [g0, g1, ...],
    "coeff"}]; value = coeff * <prod O> / <1> on the smallest cluster holding
    the sites.  Returns {(group, source, dx..., dy...): value}."""
    out = {}
    norms = {}
    for m in multisite_ops:
        src, dxs, dys = m["source"], list(m["dx"]), list(m["dy"])
        mindx, maxdx = min([0] + dxs), max([0] + dxs)
        mindy, maxdy = min([0] + dys), max([0] + dys)
        ncol, nrow = maxdx - mindx + 1, maxdy - mindy + 1
        scol, srow = -mindx, maxdy
        idx = [[lat.other(src, c - scol, srow - r) for c in range(ncol)] for r in range(nrow)]
        Tn_ = [[Tn[idx[r][c]] for c in range(ncol)] for r in range(nrow)]
        ops_ = [[np.eye(Tn[idx[r][c]].shape[4]) for c in range(ncol)] for r in range(nrow)]
        C, eTt, eTr, eTb, eTl = _cluster_args(env, idx, nrow, ncol)
        key = (idx[nrow - 1][0], nrow - 1, ncol - 1)
        if key not in norms:
            norms[key] = K.Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, Tn_, ops_)
        ops_[srow][scol] = site_ops[(src, m["ops"][0])]
        for i, (dx, dy) in enumerate(zip(dxs, dys)):
            ops_[srow - dy][scol + dx] = site_ops[(lat.other(src, dx, dy), m["ops"][i + 1])]
        val = K.Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, Tn_, ops_)
        out[(m["group"], src, *[v for pair in zip(dxs, dys) for v in pair])] = m.get("coeff", 1.0) * val / norms[key]
    return out


def measure_correlation(env, Tn, lat, site_ops, r_max, pairs, K=K):
    """correlation_function.cpp:44-220 (iTPS, CTM environment): for every
    left site and left operator, <O_left(0) O_right(r)> / <1 1> along +x
    (dx = r) and +y (dy = r) for r = 1..r_max.  site_ops[(site, group)] is
    the d x d operator; pairs = [(left_group, right_group)].  Returns
    {(left_group, left_site, right_group, dx, dy): value}."""
    out = {}
    r_ops = {}
    for lg, rg in pairs:
        r_ops.setdefault(lg, []).append(rg)
    rot = (3, 0, 1, 2, 4) if Tn[0].ndim == 5 else (3, 0, 1, 2, 4, 5)
    for left in range(lat.N_UNIT):
        eye_l = np.eye(Tn[left].shape[4])
        for lg, rgs in r_ops.items():
            if (left, lg) not in site_ops:
                continue
            op = site_ops[(left, lg)]
            # horizontal
            A = K.StartCorrelation_iTPS_CTM(env.C1[left], env.C4[left], env.eTt[left], env.eTb[left],
                                            env.eTl[left], Tn[left], op)
            An = K.StartCorrelation_iTPS_CTM(env.C1[left], env.C4[left], env.eTt[left], env.eTb[left],
                                             env.eTl[left], Tn[left], eye_l)
            right = left
            for r in range(r_max):
                right = lat.right(right)
                args = (env.C2[right], env.C3[right], env.eTt[right], env.eTr[right], env.eTb[right], Tn[right])
                norm = K.FinishCorrelation_iTPS_CTM(An, *args, np.eye(Tn[right].shape[4]))
                for rg in rgs:
                    if (right, rg) in site_ops:
                        out[(lg, left, rg, r + 1, 0)] = K.FinishCorrelation_iTPS_CTM(
                            A, *args, site_ops[(right, rg)]) / norm
                A = K.Transfer_iTPS_CTM(A, env.eTt[right], env.eTb[right], Tn[right])
                An = K.Transfer_iTPS_CTM(An, env.eTt[right], env.eTb[right], Tn[right])
            # vertical: the column seen as a row (site tensors rotated)
            t = Tn[left].transpose(rot)
            A = K.StartCorrelation_iTPS_CTM(env.C4[left], env.C3[left], env.eTl[left], env.eTr[left],
                                            env.eTb[left], t, op)
            An = K.StartCorrelation_iTPS_CTM(env.C4[left], env.C3[left], env.eTl[left], env.eTr[left],
                                             env.eTb[left], t, eye_l)
            right = left
            for r in range(r_max):
                right = lat.top(right)
                t = Tn[right].transpose(rot)
                args = (env.C1[right], env.C2[right], env.eTl[right], env.eTt[right], env.eTr[right], t)
                norm = K.FinishCorrelation_iTPS_CTM(An, *args, np.eye(Tn[right].shape[4]))
                for rg in rgs:
                    if (right, rg) in site_ops:
                        out[(lg, left, rg, 0, r + 1)] = K.FinishCorrelation_iTPS_CTM(
                            A, *args, site_ops[(right, rg)]) / norm
                A = K.Transfer_iTPS_CTM(A, env.eTl[right], env.eTr[right], t)
                An = K.Transfer_iTPS_CTM(An, env.eTl[right], env.eTr[right], t)
    return out


def correlation_lengths(env, Tn, lat, prm: TM.TMParameters):
    """correlation_length.cpp:24-55, 97-124: [(dir, coord, xi, [e_i])]."""
    tm = TM.TransferMatrixCTM(lat, Tn, env)
    out = []
    for d in (0, 1):
        W = lat.LY if d == 0 else lat.LX
        L = lat.LX_noskew if d == 0 else lat.LY_noskew
        for fixed in range(W):
            ev = tm.eigenvalues(d, fixed, prm)
            xi, es = TM.correlation_length(ev, L)
            out.append((d, fixed, xi, es, ev))
    return out
