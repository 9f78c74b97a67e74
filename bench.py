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
check of every sweep (epsilon = 1e-300: checked, never reached).

Passes, in order:
  1. timed region (the `value`): no per-launch instrumentation;
  2. the same number of sweeps again with HIP events around every GEMM
     launch (recorded on the launch's own stream): per-kernel report on
     stderr and the `timed_region_*` roofline fields;
  3. one sweep with the projector chains serialised on one stream: the
     dominant kernel (the GEMM instantiation and shape carrying the most
     flops, R = LT*LB, 2500^3) runs alone on the device, as under rocprofv3's
     kernel trace; its average launch time gives `roofline.achieved`;
  4. accuracy: a fresh environment run for the oracle's sweep count, energy,
     staggered magnetisation and the correlation lengths against the
     oracle's FP64 (LAPACK SVD) run of the same problem;
  5. mixed precision (contractions of the moves on the FP32 MFMA, modes 1
     and 2, or on the bf16 MFMA with split hi/lo operands, mode 3): the same
     timed region and accuracy check;
  6. the CPU baseline: the numpy oracle (same contraction order, LAPACK SVD)
     running one full sweep (8 moves + convergence check) on the GPU's
     environment on this host's cores.

    python bench.py [--steps K --warmup W]     # the driver's command
    python bench.py --probe-gemm               # workload for rocprofv3 --pmc
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
# algorithmic flop of one sweep at D=5, chi=100 in the reference's
# contraction order (projectors with the full R = LT*LB, 16 Calc_Next_CTM,
# 32 Calc_Next_eT), 2 flop per multiply-add
SWEEP_FLOP_D5_CHI100 = 7.448e11


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


def host_cpu():
    """(model name, logical CPUs available to this process, physical cores)."""
    model, phys = "unknown", set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if "core id" in cur:
                    phys.add((cur.get("physical id", "0"), cur["core id"]))
                cur = {}
                continue
            k, v = [x.strip() for x in line.split(":", 1)]
            cur[k] = v
            if k == "model name" and model == "unknown":
                model = v
        if "core id" in cur:
            phys.add((cur.get("physical id", "0"), cur["core id"]))
    except OSError:
        pass
    logical = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, logical, len(phys) or None


def blas_threads():
    try:
        from threadpoolctl import threadpool_info

        return max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:  # pragma: no cover
        return os.cpu_count() or 1


def cpu_baseline(env, chi):
    """One full sweep of the oracle (numpy/LAPACK, FP64: 2 left, 2 right, 2
    top, 2 bottom moves with 16 LAPACK SVDs of 2500 x 2500, then the
    convergence check) on the GPU's current environment; returns
    (sweeps/s, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import ctm as O
    from oracle.lattice import SquareLattice

    lat = SquareLattice(2, 2)
    oenv = O.Env(4)
    for k in ("C1", "C2", "C3", "C4", "eTt", "eTr", "eTb", "eTl"):
        for i in range(4):
            getattr(oenv, k)[i] = env.get(k, i).numpy()
    Tn = [env.get("Tn", i).numpy() for i in range(4)]
    p = O.PEPSParameters(CHI=chi, CTM_Convergence_Epsilon=0.0)
    old = {n: list(getattr(oenv, n)) for n in ("C1", "C2", "C3", "C4")}
    t0 = time.perf_counter()
    O.sweep(oenv, Tn, p, lat)
    O.Check_Convergence_CTM(oenv, old, p, lat)
    dt = time.perf_counter() - t0
    return 1.0 / dt, dt


def accuracy(ta, ctx, Tn, H, fx, kw, chi, tag):
    """Fresh environment, the oracle's sweep count: relative errors of the
    energy, the staggered magnetisation and the worst correlation length
    (10 leading transfer-matrix eigenvalues) against the oracle's FP64 run."""
    chk = ta.Environment(ctx, 2, 2)
    chk.set_site_tensors(Tn)
    chk.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=fx["sweeps"], **kw), initialize=True)
    E = energy(chk, H)
    e_err = abs(E - fx["energy_per_site"]) / abs(fx["energy_per_site"])
    ms = staggered_magnetization(chk)
    ms_err = abs(ms - fx["staggered_magnetization"]) / fx["staggered_magnetization"]
    xi_err = 0.0
    for c in fx["correlation"]:
        ev = chk.transfer_matrix_eigenvalues(c["dir"], c["fixed"], ta.tm_parameters(num_eigvals=10))
        xi = -2.0 / np.log(abs(ev[1]) / abs(ev[0]))
        xi_err = max(xi_err, abs(xi - c["xi"]) / c["xi"])
    log(f"{tag}: after {fx['sweeps']} sweeps E = {E:.15f} (oracle {fx['energy_per_site']:.15f}), "
        f"rel err energy {e_err:.2e}, m_s {ms_err:.2e}, worst xi {xi_err:.2e} "
        f"(10 leading eigenvalues, chi^2 = {chi ** 2})")
    return e_err, ms_err, xi_err


def timed_sweeps(ctx, env, ta, kw, steps):
    ctx.synchronize()
    t0 = time.perf_counter()
    n, _ = env.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=steps, **kw), initialize=False)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    assert n == steps
    return dt


def probe_gemm():
    """Workload for counter collection (rocprofv3 --pmc): ten launches of the
    dominant CTMRG GEMM shape (R = LT*LB at D=5, chi=100: 2500 x 2500 x 2500
    FP64, row-major operands), five Calc_Next_eT at D=5, chi=100 (each one
    fused double-layer site kernel, edge layout: 50 MB read, 50 MB written,
    besides its two GEMMs), one left-block projector at D=5, chi=100 (two
    fused site kernels, corner layout, same bytes each), and five launches
    of a plain copy kernel of known size (4096 x 4096 FP64 transpose: 128 MiB
    read, 128 MiB written) whose counter readings calibrate FETCH_SIZE /
    WRITE_SIZE."""
    import tenes_amd as ta

    ctx = ta.Context(0)
    rng = np.random.default_rng(0)
    n = 2500
    A = ctx.tensor(rng.standard_normal((n, n)))
    B = ctx.tensor(rng.standard_normal((n, n)))
    for _ in range(10):
        ta.tensordot(A, B, [1], [0])
    ctx.synchronize()
    chi, D = 100, 5
    t = ctx.tensor
    eT = t(rng.standard_normal((chi, chi, D, D)))
    Tn = t(rng.standard_normal((D, D, D, D, 2)))
    PU = t(rng.standard_normal((chi, D, D, chi)))
    PL = t(rng.standard_normal((chi, D, D, chi)))
    for _ in range(5):
        ta.calc_next_et(eT, Tn, PU, PL)
    C = [t(rng.standard_normal((chi, chi))) for _ in range(2)]
    E = [t(rng.standard_normal((chi, chi, D, D))) for _ in range(4)]
    ta.calc_projector_left_block(C[0], C[1], E[0], E[1], E[2], E[3], Tn, Tn, ta.peps_parameters(CHI=chi))
    ctx.synchronize()
    X = ctx.tensor(rng.standard_normal((4096, 4096)))
    for _ in range(5):
        ta.transpose(X, [1, 0])
    ctx.synchronize()
    print("probe: 10 GEMM 2500^3, 5 Calc_Next_eT, 1 left-block projector, 5 copies of 128 MiB done",
          flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed CTM sweeps")
    ap.add_argument("--warmup", type=int, default=3, help="untimed CTM sweeps")
    ap.add_argument("--D", type=int, default=5)
    ap.add_argument("--chi", type=int, default=100)
    ap.add_argument("--su-steps", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the untimed accuracy runs against the oracle fixture")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision passes")
    ap.add_argument("--power-iters", type=int, default=None)
    ap.add_argument("--autotune", type=int, default=0, choices=(0, 1),
                    help="GEMM autotuning of the context (tuned during the warm-up sweeps)")
    ap.add_argument("--warm-iters", type=int, default=None,
                    help="range-finder iterations of a warm-started projector SVD (library default 8)")
    ap.add_argument("--mixed", type=int, default=0, choices=(0, 1, 2, 3),
                    help="headline in mixed precision: contractions of the moves on the FP32 MFMA, 1 all but "
                         "the projector matrix R = LT LB, 2 R as well; 3 all on the bf16 MFMA with split hi/lo "
                         "operands (projector SVDs and the rest FP64)")
    ap.add_argument("--host-cut", action="store_true",
                    help="decide the projector cut on the host (one synchronisation per projector, the "
                         "reference's shapes) instead of on the device")
    ap.add_argument("--graph", type=int, default=0, choices=(0, 1),
                    help="capture a sweep into a hipGraph and replay it (the projector cut on the device)")
    ap.add_argument("--moves-in-flight", type=int, default=None,
                    help="bound on CTM moves issued by the host but unfinished on the device (0: none)")
    ap.add_argument("--probe-gemm", action="store_true",
                    help="run only the GEMM + copy workload for rocprofv3 counter passes")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (the CTMRG hot path of one problem runs on a single "
              f"MI355X here); --gpus {args.gpus} / WORLD_SIZE {world} is not supported.", flush=True)
        sys.exit(2)
    if args.probe_gemm:
        probe_gemm()
        return

    import tenes_amd as ta

    ctx = ta.Context(0, seed=2024)
    if args.autotune:
        ctx.set_autotune(True)
    if args.moves_in_flight is not None:
        ctx.set_max_moves_in_flight(args.moves_in_flight)
    fx = load_fixture(args.D, args.chi, args.su_steps)
    env, H, Tn = build_inputs(ta, ctx, args.D, args.su_steps, 0.01, fx)
    # a sweep is TeNeS's loop body: the moves and Check_Convergence_CTM.  An
    # epsilon of 1e-300 makes every sweep take the convergence check (the
    # library skips the checks a run with epsilon <= 0 cannot observe) and
    # never stops early (sig_max stays far above it)
    kw = dict(CHI=args.chi, CTM_Convergence_Epsilon=1e-300)
    if args.host_cut:
        kw["projector_cut_on_device"] = 0
    if args.graph:
        kw["graph_sweeps"] = 1
    if args.power_iters is not None:
        kw["svd_power_iters"] = args.power_iters
    if args.warm_iters is not None:
        kw["svd_warm_power_iters"] = args.warm_iters
    kw_head = dict(kw, contraction_fp32=args.mixed) if args.mixed else kw
    t0 = time.perf_counter()
    env.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=args.warmup, **kw_head), initialize=True)
    ctx.synchronize()
    log(f"warmup: {args.warmup} sweeps in {time.perf_counter() - t0:.2f} s")

    # 1. timed region, no instrumentation
    ctx.reset_stats()
    dt = timed_sweeps(ctx, env, ta, kw_head, args.steps)
    flops, ncalls = ctx.stats()
    nsvd, nrounds = ctx.svd_stats()
    sweeps_per_s = args.steps / dt
    if args.graph:
        log(f"sweep graphs (captures, replays) so far: {ctx.graph_stats()}")
    log(f"timed: {args.steps} sweeps in {dt:.3f} s -> {sweeps_per_s:.3f} sweeps/s "
        f"({SWEEP_FLOP_D5_CHI100 * sweeps_per_s / 1e12:.2f} TFLOP/s algorithmic); {ncalls} GEMMs "
        f"({flops / dt / 1e12:.2f} TFLOP/s issued); {nsvd} truncated SVDs, "
        f"{nrounds / max(nsvd, 1):.0f} Jacobi rounds each")

    # 2. the same sweeps with per-launch events (GEMM report)
    ctx.reset_kernel_report()
    ctx.set_profiling(True)
    dt_ins = timed_sweeps(ctx, env, ta, kw_head, args.steps)
    ctx.set_profiling(False)
    rep = ctx.kernel_report()
    dom_name, (dom_n, dom_ms, dom_fl) = max(rep.items(), key=lambda kv: kv[1][2])
    log(f"instrumented pass: {args.steps} sweeps in {dt_ins:.3f} s; GEMMs by device time:")
    for k, v in sorted(rep.items(), key=lambda kv: -kv[1][1])[:12]:
        log(f"  {k}: {v[0]} launches, {v[1]:.2f} ms, {v[2] / (v[1] * 1e-3) / 1e12 if v[1] else 0:.2f} TFLOP/s")
    timed_achieved = dom_fl / (dom_ms * 1e-3) / 1e12
    E = energy(env, H)

    # 3. one sweep with the projector chains serialised: the dominant GEMM
    # alone on the device, as under rocprofv3's kernel trace
    ctx.set_max_lanes(1)
    ctx.reset_kernel_report()
    ctx.set_profiling(True)
    env.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=1, **kw_head), initialize=False)
    ctx.synchronize()
    ctx.set_profiling(False)
    ctx.set_max_lanes(8)
    ser_n, ser_ms, ser_fl = ctx.kernel_report().get(dom_name, (0, 0.0, 0.0))
    ser_achieved = ser_fl / (ser_ms * 1e-3) / 1e12 if ser_ms else None
    log(f"dominant GEMM {dom_name}: {ser_ms / max(ser_n, 1):.3f} ms per launch alone on the device "
        f"-> {ser_achieved or 0:.2f} TFLOP/s; {dom_ms / dom_n:.3f} ms between its events inside the "
        f"overlapped sweeps")

    # 4. accuracy of the headline mode
    e_err = ms_err = xi_err = None
    if fx is not None and not args.no_check:
        e_err, ms_err, xi_err = accuracy(ta, ctx, Tn, H, fx, kw_head, args.chi,
                                         f"accuracy (mode {args.mixed})" if args.mixed else "accuracy (FP64)")

    # 5. mixed precision, measured in the same run
    mixed = []
    if not args.no_mixed and not args.mixed:
        for mode in (1, 2, 3):
            kwm = dict(kw, contraction_fp32=mode)
            menv = ta.Environment(ctx, 2, 2)
            menv.set_site_tensors(Tn)
            menv.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=args.warmup, **kwm), initialize=True)
            mdt = timed_sweeps(ctx, menv, ta, kwm, args.steps)
            rec = {"mode": mode,
                   "fp32_contractions": {1: "FP32 MFMA: all but R = LT*LB",
                                         2: "FP32 MFMA: all, R = LT*LB included",
                                         3: "bf16 MFMA, split hi/lo operands (3 products, FP32 "
                                            "accumulation): all, R = LT*LB included"}[mode],
                   "sweeps_per_s": args.steps / mdt, "ms_per_step": mdt / args.steps * 1e3}
            log(f"mixed mode {mode}: {args.steps} sweeps in {mdt:.3f} s -> {args.steps / mdt:.3f} sweeps/s")
            if fx is not None and not args.no_check:
                a, b, c = accuracy(ta, ctx, Tn, H, fx, kwm, args.chi, f"accuracy (mixed mode {mode})")
                rec.update(energy_rel_err_vs_fp64=a, staggered_magnetization_rel_err_vs_fp64=b,
                           correlation_length_rel_err_vs_fp64=c)
            mixed.append(rec)
            del menv

    traffic = load_traffic(dom_name)
    peak = FP32_MFMA_PEAK_TFLOPS if dom_name.startswith("gemm_f32") else FP64_MFMA_PEAK_TFLOPS

    # 6. CPU baseline
    cpu = None
    model, logical, phys = host_cpu()
    if not args.no_cpu:
        cps, cdt = cpu_baseline(env, args.chi)
        threads = blas_threads()
        log(f"cpu baseline: 1 oracle sweep in {cdt:.2f} s on {threads} BLAS threads ({model}; "
            f"{logical} logical CPUs, {phys} physical cores) -> {cps:.5f} sweeps/s")
        cpu = {"value": cps, "unit": "sweeps/s", "cores": threads, "kind": "port",
               "sample": "1 full sweep of the numpy oracle (2 left, 2 right, 2 top, 2 bottom moves: 16 "
                         "projectors with a 2500 x 2500 LAPACK SVD, 16 Calc_Next_CTM, 32 Calc_Next_eT; then "
                         "the convergence check) on the GPU's environment",
               "cpu_model": model, "logical_cpus": logical, "physical_cores": phys}

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
        # no published number exists for this metric; the CPU
        # ratio is reported separately below
        "vs_baseline": None,
        "dtype": (f"fp32-mfma contractions (mode {args.mixed}), f64 SVD/QR/Arnoldi" if args.mixed else "f64"),
        "data": ("synthetic: seeded simple-update Heisenberg site tensors (oracle run stored in "
                 "tests/golden/heisenberg_d5_chi100.json)" if fx is not None else
                 "synthetic: seeded simple-update Heisenberg site tensors generated on the GPU"),
        "config": {"workload": "CTMRG sweep, Heisenberg AFM 2x2 unit cell, D=5, d=2, chi=100, "
                               + ("mixed precision" if args.mixed else "FP64"),
                   "D": args.D, "chi": args.chi, "unit_cell": [2, 2], "su_steps": args.su_steps,
                   "projector": "corner", "svd": "GPU truncated FP64 SVD (randomized range finder + Jacobi)"},
        "algorithmic_tflops": SWEEP_FLOP_D5_CHI100 * sweeps_per_s / 1e12 if (args.D, args.chi) == (5, 100)
        else None,
        "roofline": {"bound": "mfma", "kernel": dom_name,
                     "achieved": ser_achieved if ser_achieved else timed_achieved, "peak": peak,
                     "unit": "TFLOP/s",
                     "frac": (ser_achieved if ser_achieved else timed_achieved) / peak,
                     "traffic": traffic,
                     "launches": ser_n if ser_n else dom_n,
                     "avg_launch_ms": ser_ms / ser_n if ser_n else dom_ms / dom_n,
                     "measured": "HIP events on the launch's stream around each launch, one sweep with the "
                                 "projector chains serialised (pass 3)",
                     "timed_region_launches": dom_n,
                     "timed_region_avg_launch_ms": dom_ms / dom_n,
                     "timed_region_achieved": timed_achieved,
                     "timed_region_frac": timed_achieved / peak},
        "cpu_baseline": cpu,
        "speedup_vs_cpu_baseline": sweeps_per_s / cpu["value"] if cpu else None,
        "energy_per_site_fp64": E,
        "energy_rel_err_vs_fp64": e_err,
        "staggered_magnetization_rel_err_vs_fp64": ms_err,
        "correlation_length_rel_err_vs_fp64": xi_err,
        "mixed_precision": mixed,
        "gemm_tflops_issued": flops / dt / 1e12,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
