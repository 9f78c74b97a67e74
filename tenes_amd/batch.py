"""Many independent CTMRG problems at once (tenes_batch_calc_ctm_environment).

One projector chain of a CTM sweep is latency-bound (a few CUs busy); a field
sweep of independent problems fills the device when the problems are spread
over several library contexts, each with its own streams, run from one host
thread per context.  The workload here is the batched transverse-field Ising
sweep: n problems (default 48), hx_k = 3k/(n-1), Jz = -1, 2x2 unit cell, D
and chi cycling in lockstep through (3, 32), (4, 48), (5, 64), site tensors
from 300 simple-update steps (tau = 0.01) on the GPU from the ferro state.

    python -m tenes_amd.batch [--problems 48] [--contexts 1] [--sweeps 5]
                              [--graph 1] [--lanes 8]

prints one JSON line: aggregate sweeps/s over all problems (batched and, for
comparison, one context), and each problem's energy per site.  The
environments of one context advance in lockstep (one sweep each per round,
one batched convergence check); with --graph 1 each environment's sweep is
captured once into a hipGraph and the replays of one round run side by side
on the context's lanes (--lanes).  Streams only run side by side when the
process has enough hardware queues: main() raises GPU_MAX_HW_QUEUES (HIP's
default 4) to 16.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

from . import core, inputs

SHAPES = ((3, 32), (4, 48), (5, 64))


def tfi_problem(k: int, n: int):
    """(hx, D, chi) of problem k of n."""
    D, chi = SHAPES[k % len(SHAPES)]
    return 3.0 * k / max(n - 1, 1), D, chi


def _each_context(envs, fn):
    """fn(env) for every env, one host thread per context (the library calls
    release the GIL; one context is only ever used by one thread)."""
    groups = {}
    for e in envs:
        groups.setdefault(id(e.ctx), []).append(e)
    errors = []

    def work(group):
        try:
            for e in group:
                fn(e)
        except Exception as exc:  # re-raised in the caller
            errors.append(exc)

    threads = [threading.Thread(target=work, args=(g,)) for g in groups.values()]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]


def prepare(ctxs, n: int, su_steps: int = 300, tau: float = 0.01, seed: int = 11, noise: float = 0.01,
            shapes=SHAPES):
    """Environments of the n problems, problem k on context k % len(ctxs),
    with their simple-update site tensors.  Returns (envs, hamiltonians)."""
    envs, hams = [], []
    for k in range(n):
        hx = 3.0 * k / max(n - 1, 1)
        D, chi = shapes[k % len(shapes)]
        env = core.Environment(ctxs[k % len(ctxs)], 2, 2)
        env.set_site_tensors(inputs.initial_tensors(4, D, 2, [[1.0, 0.0]] * 4, noise, seed))
        env.problem = (hx, D, chi)
        H = inputs.bond_hamiltonian(Jz=-1.0, hx=hx)
        env.su_ops = [(s, leg, inputs.gate(H, tau)) for s, leg in inputs.square_bonds(2, 2)]
        envs.append(env)
        hams.append(H)
    _each_context(envs, lambda e: e.simple_update(e.su_ops, su_steps))
    return envs, hams


def energy(env, H) -> float:
    """Energy per site: the bond values of the four +x and four +y bonds
    (the field is part of the bond operator)."""
    E = 0.0
    for s, leg in inputs.square_bonds(2, 2):
        dx, dy = (1, 0) if leg == 2 else (0, 1)
        E += env.contract_two_sites_op12(s, dx, dy, H) / env.contract_cluster(s, dx, dy)
    return E / 4.0


def run_ctm(envs, sweeps: int, initialize: bool, **kw):
    """All problems for `sweeps` CTM sweeps (epsilon 0: no early exit) in one
    tenes_batch_calc_ctm_environment call; returns seconds."""
    params = [core.peps_parameters(CHI=e.problem[2], Max_CTM_Iteration=sweeps, CTM_Convergence_Epsilon=0.0, **kw)
              for e in envs]
    t0 = time.perf_counter()
    counts, _ = core.batch_calc_ctm_environment(envs, params, initialize)
    for e in envs:
        e.ctx.synchronize()
    dt = time.perf_counter() - t0
    assert all(c == sweeps for c in counts), counts
    return dt


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--problems", type=int, default=48)
    # one context with 8 lanes and captured sweeps measured fastest (38
    # sweeps/s); captured sweeps on several contexts oversubscribe the
    # hardware queues (4 contexts: 25-26)
    ap.add_argument("--contexts", type=int, default=1)
    ap.add_argument("--sweeps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--su-steps", type=int, default=300)
    ap.add_argument("--graph", type=int, default=1, choices=(0, 1),
                    help="capture each problem's sweep into a hipGraph and replay it")
    ap.add_argument("--lanes", type=int, default=8, help="side streams per context")
    ap.add_argument("--compare-one-context", type=int, default=1, choices=(0, 1),
                    help="also time the problems on one context, uncaptured")
    ap.add_argument("--mixed", type=int, default=0, choices=(0, 1, 2, 3),
                    help="also run the batch with contraction_fp32 = 1, 2 or 3 and report each "
                         "problem's energy error against the FP64 batch")
    args = ap.parse_args(argv)
    # read by the HIP runtime when the first context initialises it
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"

    def log(msg):
        print(f"[batch] {msg}", file=sys.stderr, flush=True)

    ctxs = [core.Context(0, seed=100 + i) for i in range(args.contexts)]
    for c in ctxs:
        c.set_max_lanes(args.lanes)
    t0 = time.perf_counter()
    envs, hams = prepare(ctxs, args.problems, args.su_steps)
    log(f"{args.problems} problems, simple update ({args.su_steps} steps) in {time.perf_counter() - t0:.1f} s "
        f"on {args.contexts} contexts")
    kw = dict(graph_sweeps=args.graph)
    run_ctm(envs, args.warmup, True, **kw)
    dt = run_ctm(envs, args.sweeps, False, **kw)
    rate = args.problems * args.sweeps / dt
    log(f"batched: {args.problems} x {args.sweeps} sweeps in {dt:.2f} s -> {rate:.1f} sweeps/s")
    energies = [energy(e, H) for e, H in zip(envs, hams)]
    out = {
        "metric": "aggregate CTMRG sweeps/s, batched transverse-field Ising field sweep",
        "value": rate,
        "unit": "sweeps/s",
        "problems": args.problems,
        "contexts": args.contexts,
        "lanes": args.lanes,
        "graph_sweeps": args.graph,
        "graph_stats": [c.graph_stats() for c in ctxs],
        "sweeps": args.sweeps,
        "warmup": args.warmup,
        "dtype": "f64",
        "config": {"workload": "TFI Jz=-1, hx=3k/(n-1), 2x2, (D,chi) cycling (3,32),(4,48),(5,64), "
                               f"{args.su_steps} SU steps tau=0.01"},
        "energies": energies,
    }
    if args.compare_one_context:
        # the same problems from one context, uncaptured, same sweep count
        single = [core.Environment(ctxs[0], 2, 2) for _ in envs]
        for s, e in zip(single, envs):
            s.set_site_tensors(e.site_tensors())
            s.problem = e.problem
        run_ctm(single, args.warmup, True)
        dt1 = run_ctm(single, args.sweeps, False)
        rate1 = args.problems * args.sweeps / dt1
        log(f"one context, uncaptured: {args.problems} x {args.sweeps} sweeps in {dt1:.2f} s -> "
            f"{rate1:.1f} sweeps/s")
        e1 = [energy(e, H) for e, H in zip(single, hams)]
        out["one_context_value"] = rate1
        out["max_energy_diff_batched_vs_one_context"] = float(np.max(np.abs(np.array(energies) - np.array(e1))))
        del single
    if args.mixed:
        # the same batch with the moves' contractions on the FP32 MFMA
        run_ctm(envs, args.warmup, True, contraction_fp32=args.mixed, **kw)
        dtm = run_ctm(envs, args.sweeps, False, contraction_fp32=args.mixed, **kw)
        em = [energy(e, H) for e, H in zip(envs, hams)]
        err = [abs(a - b) / abs(b) for a, b in zip(em, energies)]
        log(f"mixed mode {args.mixed}: {args.problems} x {args.sweeps} sweeps in {dtm:.2f} s -> "
            f"{args.problems * args.sweeps / dtm:.1f} sweeps/s; energy rel err vs FP64 max {max(err):.2e}")
        out["mixed"] = {"contraction_fp32": args.mixed, "value": args.problems * args.sweeps / dtm,
                        "energy_rel_err": err, "max_energy_rel_err": max(err)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
