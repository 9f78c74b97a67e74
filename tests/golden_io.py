"""Helpers that turn the committed JSON fixtures (tests/golden/*.json, made by
tests/golden/make_fixtures.py from TeNeS's own test data) into operators,
lattices and expected values."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name: str) -> dict:
    return json.loads((GOLDEN / f"{name}.json").read_text())


def tensor_from_elements(dims, elements, is_real):
    A = np.zeros(dims, dtype=np.complex128)
    r = len(dims)
    for e in elements:
        A[tuple(e[:r])] = e[r] + 1j * e[r + 1]
    return A.real.copy() if is_real else A


def unitcell(fx):
    """(LX, LY, skew, pdims, vdims, initial_dirs, noises) from [tensor]."""
    t = fx["tensor"]
    LX, LY = t["L_sub"]
    N = LX * LY
    pdims = [0] * N
    vdims = [None] * N
    dirs = [None] * N
    noises = [0.0] * N
    for uc in t["unitcell"]:
        sites = uc["index"] or list(range(N))
        for s in sites:
            pdims[s] = uc["physical_dim"]
            vd = uc["virtual_dim"]
            vdims[s] = vd if isinstance(vd, list) else [vd] * 4
            dirs[s] = uc.get("initial_state", [0.0])
            noises[s] = uc.get("noise", 0.0)
    return LX, LY, t.get("skew", 0), pdims, vdims, dirs, noises


def operators(fx, N, is_real):
    """onesite / twosite operator dicts in the format of oracle.itps."""
    onesite = []
    onesite_names = {}
    for o in fx["onesite"]:
        d = o["dim"]
        op = tensor_from_elements([d, d], o["elements"], is_real)
        sites = o["sites"] or list(range(N))
        onesite_names[o["group"]] = o["name"]
        for s in sites:
            onesite.append({"name": o["name"], "group": o["group"], "site": s, "op": op,
                            "coeff": o["coeff"]})
    twosite = []
    twosite_names = {}
    for o in fx["twosite"]:
        twosite_names[o["group"]] = o["name"]
        for b in o["bonds"]:
            e = {"name": o["name"], "group": o["group"], "source": b[0], "dx": b[1], "dy": b[2],
                 "coeff": o["coeff"]}
            if "elements" in o:
                d1, d2 = o["dim"]
                e["op"] = tensor_from_elements([d1, d2, d1, d2], o["elements"], is_real)
            else:
                e["ops"] = o["ops"]
            twosite.append(e)
    return onesite, onesite_names, twosite, twosite_names


def evolution(fx, is_real, key="evolution"):
    return [(o["source_site"], o["source_leg"],
             tensor_from_elements(o["dimensions"], o["elements"], is_real)) for o in fx[key]]


def density_table(rows, timed: bool):
    """{(time or None, id): complex} from density.dat rows."""
    out = {}
    for r in rows:
        f = r.split()
        if timed:
            out[(round(float(f[0]), 10), int(f[1]))] = float(f[2]) + 1j * float(f[3])
        else:
            name = r.split("=")[0].strip()
            v = r.split("=")[1].split()
            out[name] = float(v[0]) + 1j * float(v[1])
    return out


def pipeline_site_tensors(name: str):
    """A TeNeS test input run through the oracle's input pipeline: initial
    tensors (seed 11), the simple-update steps, then the full-update steps
    (fast full update).  Returns (fixture, lattice, pdims, Tn)."""
    from oracle import ctm, full_update as FU, simple_update as SU
    from oracle.lattice import SquareLattice

    fx = load(name)
    LX, LY, skew, pdims, vdims, dirs, noises = unitcell(fx)
    lat = SquareLattice(LX, LY, skew)
    Tn = SU.initial_tensors(lat, pdims, vdims, dirs, noises, seed=11, is_real=True)
    lam = [[np.ones(vdims[i][l]) for l in range(4)] for i in range(lat.N_UNIT)]
    Tn, lam = SU.simple_update_steps(Tn, lam, lat, evolution(fx, True), fx["parameter"]["simple_update"]["num_step"])
    p = ctm_parameters(fx)
    Tn = FU.full_update(Tn, lat, evolution(fx, True, "evolution_full"), fx["parameter"]["full_update"]["num_step"],
                        p, FU.FUParameters())
    return fx, lat, pdims, Tn


def ctm_parameters(fx):
    """PEPSParameters of the oracle from [parameter.ctm] (dimension,
    iteration_max)."""
    from oracle import ctm

    c = fx["parameter"]["ctm"]
    kw = {"CHI": c["dimension"]}
    if "iteration_max" in c:
        kw["Max_CTM_Iteration"] = c["iteration_max"]
    return ctm.PEPSParameters(**kw)


def afh_real_site_tensors():
    """AntiferroHeisenberg_real: initial tensors (seed 11), 100 simple-update
    steps and 1 full-update step (fast full update) through the oracle, as the
    TeNeS test run does before measuring.  Returns (fixture, lattice, Tn)."""
    from oracle import ctm, full_update as FU, simple_update as SU
    from oracle.lattice import SquareLattice

    fx = load("AntiferroHeisenberg_real")
    LX, LY, skew, pdims, vdims, dirs, noises = unitcell(fx)
    lat = SquareLattice(LX, LY, skew)
    Tn = SU.initial_tensors(lat, pdims, vdims, dirs, noises, seed=11, is_real=True)
    lam = [[np.ones(vdims[i][l]) for l in range(4)] for i in range(lat.N_UNIT)]
    Tn, lam = SU.simple_update_steps(Tn, lam, lat, evolution(fx, True), fx["parameter"]["simple_update"]["num_step"])
    p = ctm.PEPSParameters(CHI=fx["parameter"]["ctm"]["dimension"])
    Tn = FU.full_update(Tn, lat, evolution(fx, True, "evolution_full"), fx["parameter"]["full_update"]["num_step"],
                        p, FU.FUParameters())
    return fx, lat, pdims, Tn
