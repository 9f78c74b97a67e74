#!/usr/bin/env python3
"""CTMRG sweeps/sec at D=5, chi=100 on one MI355X (FP64 path).

Workload (the metric's configuration): Heisenberg antiferromagnet
(Jx = Jy = Jz = 1) on the 2x2 square unit cell, D = 5, d = 2, chi = 100.
Site tensors: 500 simple-update steps (tau = 0.01) from the Neel state with
0.01 noise (std::mt19937 seed 11, TeNeS's fill order) -- the oracle's
tensors stored in tests/golden/heisenberg_d5_chi100.json when present
(else the same recipe run on the GPU).  The environment is initialised from
the site tensors and `--warmup` sweeps are run untimed (chi reaches 100 in
the first move); then `--steps` sweeps are timed, each = 2 left + 2 right +
2 top + 2 bottom moves (16 projectors with a truncated FP64 SVD of a
2500 x 2500 matrix, 16 Calc_Next_CTM, 32 Calc_Next_eT) + the convergence
check (epsilon = 0: no early exit).

Reported: sweeps/s; the dominant kernel (the GEMM instantiation and shape
carrying the most flops: R = LT*LB, 2500^3) with its achieved FP64 MFMA
TFLOP/s from HIP events recorded around every launch (roofline): alone on
the device in one extra sweep with the projector chains serialised (the
figure rocprofv3's kernel trace reports), and inside the timed region, where
four chains share the device (timed_region_*); its HBM traffic per launch from the committed rocprofv3 PMC
summary (profiles/); the relative energy error against the oracle's FP64
(LAPACK SVD) run of the same problem and sweep count, from a separate untimed
run; the CPU baseline = the numpy oracle (same contraction order, LAPACK
SVD) timing one Left_move on the same environment on this host, scaled to
sweeps/s (a sweep = 8 moves of equal cost).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# A sweep runs its left and right (top and bottom) move chains side by side,
# four projector chains on four streams next to the main one; HIP's default
# of 4 hardware queues makes two of them share a queue (14.5 against 19.1
# sweeps/s measured with 16).  Read by the runtime when it initialises, so
# set before the library is loaded; a smaller value in the environment
# (HIP's default is often exported explicitly) is raised.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = "CTMRG sweeps/sec at D=5 chi=100 (MFMA TFLOP/s) + rel energy err vs FP64"
FP64_MFMA_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


FIXTURE = os.path.join(ROOT, "tests", "golden", "heisenberg_d5_chi100.json")


def load_fixture(D, chi, su_steps):
    """The oracle's run of this problem, if stored and of the same shape."""
    if not os.path.exists(FIXTURE):
        return None
    fx = json.load(open(FIXTURE))
    if (fx["D"], fx["chi"], fx["su_steps"]) != (D, chi, su_steps):
        return None
    return fx


def build_inputs(ta, ctx, D, su_steps, tau, fx, seed=11, noise=0.01):
    from tenes_amd import inputs

    LX = LY = 2
    env = ta.Environment(ctx, LX, LY)
    H = inputs.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    if fx is not None:
        Tn = [np.array(t).reshape(D, D, D, D, 2) for t in fx["site_tensors"]]
        env.set_site_tensors(Tn)
        log(f"site tensors: oracle simple update ({su_steps} steps) from {os.path.relpath(FIXTURE, ROOT)}")
        return env, H, Tn
    Tn = inputs.initial_tensors(LX * LY, D, 2, inputs.neel_dirs(LX, LY), noise, seed)
    env.set_site_tensors(Tn)
    g = inputs.gate(H, tau)
    ops = [(s, leg, g) for s, leg in inputs.square_bonds(LX, LY)]
    t0 = time.perf_counter()
    env.simple_update(ops, su_steps)
    ctx.synchronize()
    log(f"simple update: {su_steps} steps D={D} on the GPU in {time.perf_counter() - t0:.2f} s")
    return env, H, [env.get("Tn", i).numpy() for i in range(4)]


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` (name[MxNxK]) from the committed PMC
    summary profiles/gemm_traffic.txt: lines "kernel fetch_bytes write_bytes"
    already scaled by the copy-kernel calibration; None when not measured."""
    path = os.path.join(ROOT, "profiles", "gemm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == kernel:
            return float(f[1]) + float(f[2])
    return None


def staggered_magnetization(env):
    """|(1/N) sum_i parity(i) <S_i>| on the 2x2 cell (real states: S^y = 0)."""
    from tenes_amd import inputs

    par = [1, -1, -1, 1]
    sz = sx = 0.0
    for i in range(4):
        nrm = env.contract_one_site(i)
        sz += par[i] * env.contract_one_site(i, inputs.SZ) / nrm
        sx += par[i] * env.contract_one_site(i, inputs.SX) / nrm
    return float(np.hypot(sz, sx)) / 4.0


def energy(env, H):
    from tenes_amd import inputs

    E = 0.0
    for s, leg in inputs.square_bonds(2, 2):
        dx, dy = (1, 0) if leg == 2 else (0, 1)
        E += env.contract_two_sites_op12(s, dx, dy, H) / env.contract_cluster(s, dx, dy)
    return E / 4.0


def cpu_baseline(env, chi, n_moves):
    """Oracle (numpy/LAPACK, FP64) Left_move on the GPU's current
    environment; returns (sweeps/s, seconds, threads)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import ctm as O
    from oracle.lattice import SquareLattice

    try:
        from threadpoolctl import threadpool_info

        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:  # pragma: no cover
        threads = os.cpu_count() or 1
    lat = SquareLattice(2, 2)
    oenv = O.Env(4)
    for k in ("C1", "C2", "C3", "C4", "eTt", "eTr", "eTb", "eTl"):
        for i in range(4):
            getattr(oenv, k)[i] = env.get(k, i).numpy()
    Tn = [env.get("Tn", i).numpy() for i in range(4)]
    p = O.PEPSParameters(CHI=chi)
    t0 = time.perf_counter()
    for m in range(n_moves):
        O.Left_move(oenv, Tn, m % 2, p, lat)
    dt = time.perf_counter() - t0
    per_move = dt / n_moves
    return 1.0 / (8.0 * per_move), dt, threads


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed CTM sweeps")
    ap.add_argument("--warmup", type=int, default=2, help="untimed CTM sweeps")
    ap.add_argument("--D", type=int, default=5)
    ap.add_argument("--chi", type=int, default=100)
    ap.add_argument("--su-steps", type=int, default=500)
    ap.add_argument("--cpu-moves", type=int, default=1, help="oracle moves timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the untimed accuracy run against the oracle fixture")
    ap.add_argument("--power-iters", type=int, default=None)
    ap.add_argument("--autotune", type=int, default=0, choices=(0, 1),
                    help="GEMM autotuning of the context (tuned during the warm-up sweeps)")
    ap.add_argument("--warm-iters", type=int, default=None,
                    help="range-finder iterations of a warm-started projector SVD (library default 8)")
    ap.add_argument("--mixed", type=int, default=0, choices=(0, 1, 2),
                    help="contractions of the moves on the FP32 MFMA: 1 all but the projector matrix "
                         "R = LT LB, 2 R as well (projector SVDs and the rest FP64)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (the CTMRG hot path of one problem runs on a single "
              f"MI355X here); --gpus {args.gpus} / WORLD_SIZE {world} is not supported.", flush=True)
        sys.exit(2)

    import tenes_amd as ta

    ctx = ta.Context(0, seed=2024)
    if args.autotune:
        ctx.set_autotune(True)
    fx = load_fixture(args.D, args.chi, args.su_steps)
    env, H, Tn = build_inputs(ta, ctx, args.D, args.su_steps, 0.01, fx)
    kw = dict(CHI=args.chi, CTM_Convergence_Epsilon=0.0)
    if args.power_iters is not None:
        kw["svd_power_iters"] = args.power_iters
    if args.warm_iters is not None:
        kw["svd_warm_power_iters"] = args.warm_iters
    if args.mixed:
        kw["contraction_fp32"] = args.mixed
    t0 = time.perf_counter()
    env.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=args.warmup, **kw), initialize=True)
    ctx.synchronize()
    log(f"warmup: {args.warmup} sweeps in {time.perf_counter() - t0:.2f} s")

    ctx.reset_stats()
    ctx.reset_kernel_report()
    ctx.set_profiling(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    n, _ = env.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=args.steps, **kw), initialize=False)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.set_profiling(False)
    assert n == args.steps
    flops, ncalls = ctx.stats()
    nsvd, nrounds = ctx.svd_stats()
    rep = ctx.kernel_report()
    sweeps_per_s = args.steps / dt
    log(f"timed: {args.steps} sweeps in {dt:.3f} s -> {sweeps_per_s:.3f} sweeps/s; "
        f"{ncalls} GEMMs, {flops / dt / 1e12:.2f} TFLOP/s overall; {nsvd} truncated SVDs, "
        f"{nrounds / max(nsvd, 1):.0f} Jacobi rounds each")
    # dominant kernel: the GEMM (instantiation and shape) carrying the most
    # algorithmic work of the timed region
    dom_name, (dom_n, dom_ms, dom_fl) = max(rep.items(), key=lambda kv: kv[1][2])
    gemm_ms = sum(v[1] for v in rep.values())
    for k, v in sorted(rep.items(), key=lambda kv: -kv[1][1])[:12]:
        log(f"  {k}: {v[0]} launches, {v[1]:.2f} ms, {v[2] / (v[1] * 1e-3) / 1e12 if v[1] else 0:.2f} TFLOP/s")
    achieved = dom_fl / (dom_ms * 1e-3) / 1e12
    E = energy(env, H)
    log(f"energy per site (FP64): {E:.15f}; GEMM share of timed region {gemm_ms / (dt * 1e3):.1%}")
    # one more sweep with the projector lanes serialised (one stream): the
    # dominant GEMM then runs alone on the device, as under rocprofv3's kernel
    # trace, whose average launch time this one matches
    ctx.set_max_lanes(1)
    ctx.reset_kernel_report()
    ctx.set_profiling(True)
    env.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=1, **kw), initialize=False)
    ctx.synchronize()
    ctx.set_profiling(False)
    ctx.set_max_lanes(8)
    ser_n, ser_ms, ser_fl = ctx.kernel_report().get(dom_name, (0, 0.0, 0.0))
    ser_achieved = ser_fl / (ser_ms * 1e-3) / 1e12 if ser_ms else None
    log(f"dominant GEMM {dom_name}: {dom_ms / dom_n:.3f} ms per launch in the timed region (four lanes), "
        f"{ser_ms / max(ser_n, 1):.3f} ms serialised -> {ser_achieved or 0:.2f} TFLOP/s")
    # accuracy against the oracle's FP64 (LAPACK SVD) run of the same problem:
    # a fresh environment, the oracle's sweep count, untimed
    e_err = ms_err = xi_err = None
    if fx is not None and not args.no_check:
        chk = ta.Environment(ctx, 2, 2)
        chk.set_site_tensors(Tn)
        chk.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=fx["sweeps"], **kw), initialize=True)
        E_chk = energy(chk, H)
        e_err = abs(E_chk - fx["energy_per_site"]) / abs(fx["energy_per_site"])
        log(f"energy after {fx['sweeps']} sweeps: GPU {E_chk:.15f}, oracle {fx['energy_per_site']:.15f}, "
            f"rel err {e_err:.2e}")
        ms = staggered_magnetization(chk)
        ms_err = abs(ms - fx["staggered_magnetization"]) / fx["staggered_magnetization"]
        # every correlation length of the fixture (both directions, both
        # rows/columns); the worst relative error is reported
        xi_err = 0.0
        for c in fx["correlation"]:
            ev = chk.transfer_matrix_eigenvalues(c["dir"], c["fixed"], ta.tm_parameters(num_eigvals=10))
            xi = -2.0 / np.log(abs(ev[1]) / abs(ev[0]))
            err = abs(xi - c["xi"]) / c["xi"]
            xi_err = max(xi_err, err)
            log(f"correlation length dir {c['dir']} row/col {c['fixed']}: {xi:.12f} (rel err {err:.2e}, "
                f"10 leading eigenvalues, chi^2 = {args.chi ** 2})")
        log(f"staggered magnetisation {ms:.12f} (rel err {ms_err:.2e})")
    traffic = load_traffic(dom_name)
    # the dominant GEMM's matrix peak: FP32 MFMA in the mixed mode
    peak = FP32_MFMA_PEAK_TFLOPS if dom_name.startswith("gemm_f32") else FP64_MFMA_PEAK_TFLOPS

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cps, cdt, threads = cpu_baseline(env, args.chi, args.cpu_moves)
        log(f"cpu baseline: {args.cpu_moves} Left_move in {cdt:.2f} s on {threads} threads -> {cps:.5f} sweeps/s")
        cpu = {"value": cps, "unit": "sweeps/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_moves} Left_move (2 projectors with a 2500x2500 LAPACK SVD, 2 Calc_Next_CTM, "
                         f"4 Calc_Next_eT) of the numpy oracle on the same environment; sweeps/s = 1/(8 t_move)"}

    out = {
        "metric": METRIC,
        "value": sweeps_per_s,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (f"fp32-mfma contractions (mode {args.mixed}), f64 SVD/QR/Arnoldi" if args.mixed else "f64"),
        "data": ("synthetic: seeded simple-update Heisenberg site tensors (oracle run stored in "
                 "tests/golden/heisenberg_d5_chi100.json)" if fx is not None else
                 "synthetic: seeded simple-update Heisenberg site tensors generated on the GPU"),
        "config": {"workload": "CTMRG sweep, Heisenberg AFM 2x2 unit cell, D=5, d=2, chi=100, "
                               + ("mixed precision" if args.mixed else "FP64"),
                   "D": args.D, "chi": args.chi, "unit_cell": [2, 2], "su_steps": args.su_steps,
                   "projector": "corner", "svd": "GPU truncated FP64 SVD (randomized range finder + Jacobi)"},
        # the kernel alone on the device (the serialised sweep; what
        # rocprofv3's kernel trace reports), then the same launches inside
        # the timed region, where four projector chains share the device and
        # a launch's event-to-event time includes waiting for their CUs
        "roofline": {"bound": "mfma", "kernel": dom_name,
                     "achieved": ser_achieved if ser_achieved else achieved, "peak": peak,
                     "unit": "TFLOP/s",
                     "frac": (ser_achieved if ser_achieved else achieved) / peak,
                     "traffic": traffic,
                     "launches": ser_n if ser_n else dom_n,
                     "avg_launch_ms": ser_ms / ser_n if ser_n else dom_ms / dom_n,
                     "measured": "HIP events around each launch, one extra sweep with the projector "
                                 "chains serialised",
                     "timed_region_launches": dom_n,
                     "timed_region_avg_launch_ms": dom_ms / dom_n,
                     "timed_region_achieved": achieved,
                     "timed_region_frac": achieved / peak},
        "cpu_baseline": cpu,
        "energy_per_site_fp64": E,
        "energy_rel_err_vs_fp64": e_err,
        "staggered_magnetization_rel_err_vs_fp64": ms_err,
        "correlation_length_rel_err_vs_fp64": xi_err,
        "gemm_tflops_overall": flops / dt / 1e12,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
