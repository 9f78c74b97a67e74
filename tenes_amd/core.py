"""Thin Python objects over the C ABI (include/tenes_amd.h).

Every call goes to libtenes_amd.so; nothing here computes on the CPU beyond
argument marshalling.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

from . import _lib as L

KIND = {"C1": 0, "C2": 1, "C3": 2, "C4": 3, "eTt": 4, "eTr": 5, "eTb": 6, "eTl": 7, "Tn": 8}


class TenesError(RuntimeError):
    pass


class Context:
    """One GPU, one HIP stream (tenes_ctx)."""

    def __init__(self, device: int = 0, stream: int | None = None, seed: int | None = None):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.tenes_ctx_create(device, C.c_void_p(stream or 0), C.byref(h))
        if rc != 0:
            msg = self.lib.tenes_ctx_last_error(None)
            raise TenesError(f"tenes_ctx_create failed ({rc}): {msg.decode() if msg else ''}")
        self.h = h
        if seed is not None:
            self.check(self.lib.tenes_ctx_set_seed(self.h, C.c_uint64(seed)))

    def check(self, rc: int):
        if rc != 0:
            msg = self.lib.tenes_ctx_last_error(self.h)
            raise TenesError(f"tenes error {rc}: {msg.decode() if msg else ''}")

    def synchronize(self):
        self.check(self.lib.tenes_ctx_synchronize(self.h))

    def stats(self):
        f = C.c_double()
        n = C.c_int64()
        self.check(self.lib.tenes_ctx_stats(self.h, C.byref(f), C.byref(n)))
        return f.value, n.value

    def set_max_lanes(self, n: int):
        """Side streams a CTM move may overlap its rows on (1: none)."""
        self.check(self.lib.tenes_ctx_set_max_lanes(self.h, int(n)))

    def set_max_moves_in_flight(self, n: int):
        """Bound on CTM moves issued but unfinished on the device (0: none)."""
        self.check(self.lib.tenes_ctx_set_max_moves_in_flight(self.h, int(n)))

    def graph_stats(self):
        """(captures, replays) of CTM sweep graphs through this context."""
        c, r = C.c_int64(), C.c_int64()
        self.check(self.lib.tenes_ctx_graph_stats(self.h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def lane_state(self):
        """(forked, lane): whether work is split over side streams and the
        stream new work goes to (-1: the context stream)."""
        f, l = C.c_int(), C.c_int()
        self.check(self.lib.tenes_ctx_lane_state(self.h, C.byref(f), C.byref(l)))
        return bool(f.value), l.value

    def set_autotune(self, on: bool):
        """GEMM autotuning (tenes_ctx_set_autotune)."""
        self.check(self.lib.tenes_ctx_set_autotune(self.h, 1 if on else 0))

    def set_stream_k(self, on: bool):
        """Stream-K launches of the qualifying FP64 GEMMs (tenes_ctx_set_stream_k)."""
        self.check(self.lib.tenes_ctx_set_stream_k(self.h, 1 if on else 0))

    def set_fp32_contractions(self, on: bool):
        """Contractions from now on run on the FP32 MFMA (FP64 data)."""
        self.check(self.lib.tenes_ctx_set_fp32_contractions(self.h, 1 if on else 0))

    def set_fused_site(self, on: bool):
        """The enlarged corners' and Calc_Next_eT's pair of site contractions
        as one kernel (default) or as the reference's two contractions."""
        self.check(self.lib.tenes_ctx_set_fused_site(self.h, 1 if on else 0))

    def set_contraction_mode(self, mode: int):
        """Arithmetic of the contractions from now on: 0 FP64, 1 FP32 MFMA,
        2 bf16 MFMA with split hi/lo operands (FP32 accumulation)."""
        self.check(self.lib.tenes_ctx_set_fp32_contractions(self.h, int(mode)))

    def svd_stats(self):
        """(truncated SVDs, Jacobi rounds) since the last reset_stats()."""
        a = C.c_int64()
        b = C.c_int64()
        self.check(self.lib.tenes_ctx_svd_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def reset_stats(self):
        self.check(self.lib.tenes_ctx_reset_stats(self.h))

    def set_profiling(self, on: bool):
        self.check(self.lib.tenes_ctx_set_profiling(self.h, 1 if on else 0))

    def kernel_report(self) -> dict:
        """{kernel name: (launches, total_ms, total_flops)} of profiled GEMMs."""
        buf = C.create_string_buffer(1 << 20)
        self.check(self.lib.tenes_ctx_kernel_report(self.h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms, fl = line.rsplit(" ", 3)
            out[name] = (int(n), float(ms), float(fl))
        return out

    def reset_kernel_report(self):
        self.check(self.lib.tenes_ctx_reset_kernel_report(self.h))

    def tensor(self, a: np.ndarray) -> "Tensor":
        a = np.ascontiguousarray(a, dtype=np.float64)
        shape = np.array(a.shape, dtype=np.int64)
        h = C.c_void_p()
        self.check(self.lib.tenes_tensor_from_host(self.h, a.ndim, L.i64ptr(shape), L.dptr(a), C.byref(h)))
        return Tensor(self, h)

    def from_device(self, ptr: int, shape: Sequence[int]) -> "Tensor":
        shp = np.array(shape, dtype=np.int64)
        h = C.c_void_p()
        self.check(self.lib.tenes_tensor_from_device(self.h, len(shp), L.i64ptr(shp), C.c_void_p(ptr), C.byref(h)))
        return Tensor(self, h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.tenes_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tensor:
    """Device FP64 row-major tensor handle (tenes_tensor)."""

    def __init__(self, ctx: Context, h: C.c_void_p):
        self.ctx = ctx
        self.h = h

    @property
    def shape(self) -> tuple:
        lib = self.ctx.lib
        r = C.c_int()
        lib.tenes_tensor_rank(self.h, C.byref(r))
        s = np.zeros(max(r.value, 1), dtype=np.int64)
        lib.tenes_tensor_shape(self.h, L.i64ptr(s))
        return tuple(int(x) for x in s[: r.value])

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        self.ctx.check(self.ctx.lib.tenes_tensor_to_host(self.h, L.dptr(out)))
        return out

    def copy_to_device(self, ptr: int):
        """Copy into caller-owned device memory (a device pointer of the
        same HIP runtime) and wait for the copy."""
        self.ctx.check(self.ctx.lib.tenes_tensor_to_device(self.h, C.c_void_p(ptr)))

    def device_ptr(self) -> int:
        p = C.c_void_p()
        self.ctx.lib.tenes_tensor_device_ptr(self.h, C.byref(p))
        return p.value or 0

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.ctx.lib.tenes_tensor_destroy(self.h)
        except Exception:
            pass
        self.h = None


def batch_calc_ctm_environment(envs, params, initialize: bool = True):
    """Calc_CTM_Environment on several environments at once
    (tenes_batch_calc_ctm_environment): environments of one context run in
    lockstep (one sweep each per round, one batched convergence check per
    round, captured sweeps replayed side by side), different contexts
    concurrently.  params: one PepsParameters or
    one per environment.  Returns (sweep counts, sig_max values)."""
    envs = list(envs)
    n = len(envs)
    if isinstance(params, L.PepsParameters):
        params = [params]
    params = list(params)
    if len(params) not in (1, n):
        raise ValueError("one parameter set, or one per environment")
    lib = L.load()
    arr = (C.c_void_p * max(n, 1))(*[e.h for e in envs])
    prm = (L.PepsParameters * len(params))(*params)
    cnt = np.zeros(max(n, 1), dtype=np.int32)
    sig = np.zeros(max(n, 1), dtype=np.float64)
    rc = lib.tenes_batch_calc_ctm_environment(arr, n, prm, len(params), 1 if initialize else 0,
                                              L.iptr(cnt), L.dptr(sig))
    if rc != 0:
        msg = lib.tenes_ctx_last_error(None)
        raise TenesError(f"tenes_batch_calc_ctm_environment failed ({rc}): {msg.decode() if msg else ''}")
    return [int(c) for c in cnt[:n]], [float(s) for s in sig[:n]]


def Tensor_zeros(ctx: Context, shape: Sequence[int]) -> Tensor:
    """Zero-filled device tensor (tenes_tensor_zeros)."""
    shp = np.array(shape, dtype=np.int64)
    h = C.c_void_p()
    ctx.check(ctx.lib.tenes_tensor_zeros(ctx.h, len(shp), L.i64ptr(shp), C.byref(h)))
    return Tensor(ctx, h)


def tensordot(A: Tensor, B: Tensor, axA: Sequence[int], axB: Sequence[int],
              perm: Sequence[int] | None = None) -> Tensor:
    ctx = A.ctx
    a = np.array(axA, dtype=np.int32)
    b = np.array(axB, dtype=np.int32)
    p = np.array(perm, dtype=np.int32) if perm is not None else None
    h = C.c_void_p()
    ctx.check(ctx.lib.tenes_tensordot(A.h, B.h, len(a), L.iptr(a), L.iptr(b),
                                      L.iptr(p) if p is not None else None, C.byref(h)))
    return Tensor(ctx, h)


def transpose(T: Tensor, axes: Sequence[int]) -> Tensor:
    ctx = T.ctx
    a = np.array(axes, dtype=np.int32)
    h = C.c_void_p()
    ctx.check(ctx.lib.tenes_transpose(T.h, L.iptr(a), C.byref(h)))
    return Tensor(ctx, h)


def svd_top(A: Tensor, k: int, power_iters: int = 4):
    ctx = A.ctx
    U = C.c_void_p()
    V = C.c_void_p()
    s = np.zeros(max(k, 1), dtype=np.float64)
    kk = C.c_int()
    ctx.check(ctx.lib.tenes_svd_top(A.h, k, power_iters, C.byref(U), C.byref(V), L.dptr(s), C.byref(kk)))
    return Tensor(ctx, U), s[: kk.value].copy(), Tensor(ctx, V)


def peps_parameters(**kw) -> L.PepsParameters:
    """PEPS_Parameters (CTM fields) with TeNeS defaults, overridden by kw."""
    p = L.PepsParameters()
    L.load().tenes_peps_parameters_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


def fu_parameters(**kw) -> L.FuParameters:
    """tenes_fu_parameters with TeNeS's defaults, overridden by kw."""
    p = L.FuParameters()
    L.load().tenes_fu_parameters_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def full_update_bond(tensors: Sequence[Tensor], op12: np.ndarray, connect1: int, params=None):
    """Full_update_bond (core/full_update.cpp:395-441): tensors = C1..C4,
    eT1..eT6, Tn1, Tn2 (device Tensors).  Returns (Tn1_new, Tn2_new)."""
    ctx = tensors[0].ctx
    arr = (C.c_void_p * 12)(*[t.h for t in tensors])
    o = np.ascontiguousarray(op12, dtype=np.float64)
    a, b = C.c_void_p(), C.c_void_p()
    prm = params or fu_parameters()
    ctx.check(ctx.lib.tenes_full_update_bond(arr, L.dptr(o), connect1, C.byref(prm), C.byref(a), C.byref(b)))
    return Tensor(ctx, a), Tensor(ctx, b)


def tm_parameters(**kw) -> L.TmParameters:
    p = L.TmParameters()
    L.load().tenes_tm_parameters_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


def calc_projector_left_block(C1, C4, eT1, eT6, eT7, eT8, Tn1, Tn4, params):
    ctx = C1.ctx
    PU = C.c_void_p()
    PL = C.c_void_p()
    ctx.check(ctx.lib.tenes_calc_projector_left_block(
        C1.h, C4.h, eT1.h, eT6.h, eT7.h, eT8.h, Tn1.h, Tn4.h, C.byref(params), C.byref(PU), C.byref(PL)))
    return Tensor(ctx, PU), Tensor(ctx, PL)


def calc_projector_updown_blocks(tensors: Sequence[Tensor], params):
    ctx = tensors[0].ctx
    arr = (C.c_void_p * 16)(*[t.h for t in tensors])
    PU = C.c_void_p()
    PL = C.c_void_p()
    ctx.check(ctx.lib.tenes_calc_projector_updown_blocks(arr, C.byref(params), C.byref(PU), C.byref(PL)))
    return Tensor(ctx, PU), Tensor(ctx, PL)


def calc_next_ctm(C1, C4, eT1, eT6, PU, PL):
    ctx = C1.ctx
    a = C.c_void_p()
    b = C.c_void_p()
    ctx.check(ctx.lib.tenes_calc_next_ctm(C1.h, C4.h, eT1.h, eT6.h, PU.h, PL.h, C.byref(a), C.byref(b)))
    return Tensor(ctx, a), Tensor(ctx, b)


def calc_next_et(eT8, Tn1, PU, PL):
    ctx = eT8.ctx
    a = C.c_void_p()
    ctx.check(ctx.lib.tenes_calc_next_et(eT8.h, Tn1.h, PU.h, PL.h, C.byref(a)))
    return Tensor(ctx, a)


def simple_update_bond(Tn1: Tensor, Tn2: Tensor, lambda1, lambda2, op12: Tensor, connect1: int,
                       inverse_lambda_cut: float = 1e-12):
    """lambda1/lambda2: lists of the four bond-weight vectors."""
    ctx = Tn1.ctx
    l1 = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in lambda1]))
    l2 = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in lambda2]))
    a = C.c_void_p()
    b = C.c_void_p()
    lc = np.zeros(Tn1.shape[connect1], dtype=np.float64)
    ctx.check(ctx.lib.tenes_simple_update_bond(Tn1.h, Tn2.h, L.dptr(l1), L.dptr(l2), op12.h, connect1,
                                               inverse_lambda_cut, C.byref(a), C.byref(b), L.dptr(lc)))
    return Tensor(ctx, a), Tensor(ctx, b), lc


def create_environment_two_sites(tensors: Sequence[Tensor]) -> Tensor:
    ctx = tensors[0].ctx
    arr = (C.c_void_p * 12)(*[t.h for t in tensors])
    out = C.c_void_p()
    ctx.check(ctx.lib.tenes_create_environment_two_sites(arr, C.byref(out)))
    return Tensor(ctx, out)


class Environment:
    """The iTPS tensor arrays (Tn, C1..C4, eTt/eTr/eTb/eTl) on a square
    lattice unit cell, resident on the GPU (tenes_env)."""

    def __init__(self, ctx: Context, LX: int, LY: int, skew: int = 0):
        self.ctx = ctx
        self.LX, self.LY, self.skew = LX, LY, skew
        h = C.c_void_p()
        ctx.check(ctx.lib.tenes_env_create(ctx.h, LX, LY, skew, C.byref(h)))
        self.h = h
        self.N = LX * LY

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.ctx.lib.tenes_env_destroy(self.h)
        except Exception:
            pass
        self.h = None

    def set(self, kind: str, site: int, t):
        if isinstance(t, np.ndarray):
            t = self.ctx.tensor(t)
        self.ctx.check(self.ctx.lib.tenes_env_set(self.h, KIND[kind], site, t.h))

    def get(self, kind: str, site: int) -> Tensor:
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.tenes_env_get(self.h, KIND[kind], site, C.byref(h)))
        return Tensor(self.ctx, h)

    def set_site_tensors(self, Tn: Sequence[np.ndarray]):
        for i, t in enumerate(Tn):
            self.set("Tn", i, t)

    def calc_ctm_environment(self, params, initialize: bool = True):
        n = C.c_int()
        sig = C.c_double()
        self.ctx.check(self.ctx.lib.tenes_calc_ctm_environment(
            self.h, C.byref(params), 1 if initialize else 0, C.byref(n), C.byref(sig)))
        return n.value, sig.value

    def calc_ctm_environment_density(self, params, initialize: bool = True):
        """Single-layer CTM of an iTPO (site tensors (D, D, D, D, d, d)),
        Calc_CTM_Environment_density of ctm_single.cpp:737-1057."""
        n = C.c_int()
        sig = C.c_double()
        self.ctx.check(self.ctx.lib.tenes_calc_ctm_environment_density(
            self.h, C.byref(params), 1 if initialize else 0, C.byref(n), C.byref(sig)))
        return n.value, sig.value

    def move(self, direction: int, coord: int, params):
        self.ctx.check(self.ctx.lib.tenes_env_move(self.h, direction, coord, C.byref(params)))

    def save_corners(self):
        self.ctx.check(self.ctx.lib.tenes_env_save_corners(self.h))

    def check_convergence(self, params):
        c = C.c_int()
        s = C.c_double()
        self.ctx.check(self.ctx.lib.tenes_env_check_convergence(self.h, C.byref(params), C.byref(c), C.byref(s)))
        return bool(c.value), s.value

    def corner_spectrum(self, kind: str, site: int) -> np.ndarray:
        out = np.zeros(256, dtype=np.float64)
        n = C.c_int()
        self.ctx.check(self.ctx.lib.tenes_env_corner_spectrum(self.h, KIND[kind], site, L.dptr(out), C.byref(n)))
        return out[: n.value].copy()

    def contract_one_site(self, site: int, op: np.ndarray | None = None) -> float:
        out = C.c_double()
        o = np.ascontiguousarray(op, dtype=np.float64) if op is not None else None
        self.ctx.check(self.ctx.lib.tenes_env_contract_one_site(
            self.h, site, L.dptr(o) if o is not None else None, C.byref(out)))
        return out.value

    def contract_two_sites_op12(self, source: int, dx: int, dy: int, op12: np.ndarray) -> float:
        out = C.c_double()
        o = np.ascontiguousarray(op12, dtype=np.float64)
        self.ctx.check(self.ctx.lib.tenes_env_contract_two_sites_op12(
            self.h, source, dx, dy, L.dptr(o), C.byref(out)))
        return out.value

    def contract_cluster(self, source: int, dx: int, dy: int,
                         op_source: np.ndarray | None = None,
                         op_target: np.ndarray | None = None) -> float:
        out = C.c_double()
        a = np.ascontiguousarray(op_source, dtype=np.float64) if op_source is not None else None
        b = np.ascontiguousarray(op_target, dtype=np.float64) if op_target is not None else None
        self.ctx.check(self.ctx.lib.tenes_env_contract_cluster(
            self.h, source, dx, dy, L.dptr(a) if a is not None else None,
            L.dptr(b) if b is not None else None, C.byref(out)))
        return out.value

    def simple_update(self, ops, nsteps: int, inverse_lambda_cut: float = 1e-12):
        """ops: [(source_site, source_leg, op (d1, d2, d1, d2)), ...]"""
        src = np.array([o[0] for o in ops], dtype=np.int32)
        leg = np.array([o[1] for o in ops], dtype=np.int32)
        tens = [o[2] if isinstance(o[2], Tensor) else self.ctx.tensor(o[2]) for o in ops]
        arr = (C.c_void_p * max(len(tens), 1))(*[t.h for t in tens])
        self.ctx.check(self.ctx.lib.tenes_env_simple_update(
            self.h, len(ops), L.iptr(src), L.iptr(leg), arr, nsteps, inverse_lambda_cut))

    def full_update(self, ops, nsteps: int, ctm_params=None, fu_params=None):
        """TeNeS's full-update driver (src/iTPS/full_update.cpp:154-186):
        ops = [(source_site, source_leg, op), ...] with op (d_s, d_t, d_s', d_t')
        (output dimensions may differ: path-decomposed gates)."""
        src = np.array([o[0] for o in ops], dtype=np.int32)
        leg = np.array([o[1] for o in ops], dtype=np.int32)
        flat = np.concatenate([np.ascontiguousarray(o[2], dtype=np.float64).ravel() for o in ops]) \
            if ops else np.zeros(1)
        dims = np.array([np.shape(o[2]) for o in ops], dtype=np.int32).reshape(-1) if ops \
            else np.zeros(4, dtype=np.int32)
        pp = ctm_params or peps_parameters()
        fp = fu_params or fu_parameters()
        self.ctx.check(self.ctx.lib.tenes_env_full_update(
            self.h, len(ops), L.iptr(src), L.iptr(leg), L.iptr(dims), L.dptr(flat), nsteps, C.byref(pp),
            C.byref(fp)))

    def contract_multisite(self, source: int, dx, dy, ops) -> float:
        """<prod O_i> / <1> of one-site operators on the source and the
        sites (dx[i], dy[i]) from it (tenes_env_contract_multisite)."""
        dxa = np.ascontiguousarray(dx, dtype=np.int32)
        dya = np.ascontiguousarray(dy, dtype=np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(o, dtype=np.float64).ravel() for o in ops]))
        out = C.c_double()
        self.ctx.check(self.ctx.lib.tenes_env_contract_multisite(
            self.h, source, len(dxa), L.iptr(dxa) if len(dxa) else None, L.iptr(dya) if len(dya) else None,
            L.dptr(flat), C.byref(out)))
        return out.value

    def correlation(self, left_site: int, left_op: np.ndarray, right_ops, r_max: int,
                    direction: int = 0) -> np.ndarray:
        """<O_left(site) O_j(site + r)> / <1 1> for r = 1..r_max along +x
        (direction 0) or +y (1): array (r_max, len(right_ops))."""
        lo = np.ascontiguousarray(left_op, dtype=np.float64)
        ro = np.ascontiguousarray(np.stack([np.asarray(o, dtype=np.float64) for o in right_ops])) \
            if len(right_ops) else np.zeros(1)
        out = np.zeros(max(r_max * len(right_ops), 1), dtype=np.float64)
        self.ctx.check(self.ctx.lib.tenes_env_correlation(
            self.h, left_site, L.dptr(lo), len(right_ops), L.dptr(ro), r_max, direction, L.dptr(out)))
        return out[: r_max * len(right_ops)].reshape(r_max, len(right_ops))

    def bond_energy(self, bonds, grad=None):
        """Energy per site of nearest-neighbour bond operators with the
        environment held fixed (tenes_env_bond_energy).

        bonds: [(source, dx, dy, op12 (d_source, d_target, d_source,
        d_target))].  grad: None (energy only), True (returns host arrays
        dE/dTn_i) or a list of N_UNIT device Tensors shaped like the site
        tensors, overwritten in place.  Returns (E, grads or None)."""
        src = np.array([b[0] for b in bonds], dtype=np.int32)
        dx = np.array([b[1] for b in bonds], dtype=np.int32)
        dy = np.array([b[2] for b in bonds], dtype=np.int32)
        ops = np.concatenate([np.ascontiguousarray(b[3], dtype=np.float64).ravel() for b in bonds]) \
            if bonds else np.zeros(1)
        E = C.c_double()
        out = None
        arr = None
        if grad is not None:
            if grad is True:
                out = [Tensor_zeros(self.ctx, self.get("Tn", i).shape) for i in range(self.N)]
            else:
                out = list(grad)
            if len(out) != self.N:
                raise ValueError("one gradient tensor per site expected")
            arr = (C.c_void_p * self.N)(*[g.h for g in out])
        self.ctx.check(self.ctx.lib.tenes_env_bond_energy(
            self.h, len(bonds), L.iptr(src), L.iptr(dx), L.iptr(dy), L.dptr(ops), C.byref(E), arr))
        if grad is True:
            return E.value, [g.numpy() for g in out]
        return E.value, out

    def get_lambda(self, site: int, leg: int) -> np.ndarray:
        out = np.zeros(4096, dtype=np.float64)
        n = C.c_int()
        self.ctx.check(self.ctx.lib.tenes_env_get_lambda(self.h, site, leg, L.dptr(out), C.byref(n)))
        return out[: n.value].copy()

    def site_tensors(self) -> list:
        return [self.get("Tn", i).numpy() for i in range(self.N)]

    def save_tensors(self, directory: str, chi: int = 0) -> None:
        """Checkpoint (iTPS::save_tensors, saveload_tensors.cpp:38-86):
        params.dat, one file per tensor and the bond weights in directory."""
        self.ctx.check(self.ctx.lib.tenes_env_save_tensors(self.h, os.fsencode(directory), chi))

    def load_tensors(self, directory: str, chi: int = 0) -> None:
        """Restart from a checkpoint (iTPS::load_tensors, saveload_tensors.cpp:
        88-230), each tensor resized to the shape this environment holds for
        it; chi > 0 sets the environment legs of tensors it does not hold."""
        self.ctx.check(self.ctx.lib.tenes_env_load_tensors(self.h, os.fsencode(directory), chi))

    def transfer_matrix_eigenvalues(self, direction: int, fixed: int, params=None) -> np.ndarray:
        params = params or tm_parameters()
        out = np.zeros(2 * max(params.num_eigvals, 1), dtype=np.float64)
        n = C.c_int()
        self.ctx.check(self.ctx.lib.tenes_env_transfer_matrix_eigenvalues(
            self.h, direction, fixed, C.byref(params), L.dptr(out), C.byref(n)))
        return out[0: 2 * n.value: 2] + 1j * out[1: 2 * n.value: 2]
