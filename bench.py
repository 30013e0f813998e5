#!/usr/bin/env python3
"""Benchmark of the linear-static hot path on one MI355X.

Workload (default): CHEXA8 120x40x40 block, +-10% interior jitter, lognormal
E (median 200e3, sigma 0.3), nu=0.3, clamped x=0, PLOAD4 p=1.0 on the top
faces -- 203,401 nodes, 605,160 DOF. K is assembled on the GPU (EAS elements)
and stays in HBM; F is uploaded once. One step = one Jacobi-PCG solve from
u=0 to a true FP64 relative residual ||F-Ku||/||F|| < 1e-8, in FP32 SpMV /
Jacobi with FP64 refreshes and FP64 solution update. The plain FP64 GPU PCG
is timed on the same device matrix for comparison.

Prints one JSON line (see the README of the harness contract).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c1", "c2"])
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solve")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--profile", action="store_true",
                    help="short run for profilers: one solve per precision, few kernel reps")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(cfg):
    from vibestran_amd import decks
    if cfg == "c1":
        return decks.config1_hex_cantilever()
    return decks.config2_hex_jitter()


def cpu_baseline(model, eq, n, F, seconds):
    """Oracle FP64 Jacobi-PCG (the CPU restatement) on the host cores, timed
    for a bounded number of iterations of the same system."""
    import oracle as O
    nthreads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    t0 = time.perf_counter()
    K = O.assemble_solids(n, model.elem_type, model.elem_ptr, model.conn, model.xyz, eq, model.E,
                          model.nu)
    t_asm = time.perf_counter() - t0
    _, s = O.pcg_jacobi_fixed(K, F, 5, nthreads)
    per_it = max(s / 5, 1e-6)
    iters = int(max(5, min(5000, seconds / per_it)))
    _, s = O.pcg_jacobi_fixed(K, F, iters, nthreads)
    return {"value": iters / s, "unit": "PCG iters/s", "cores": nthreads, "kind": "port",
            "sample": f"{iters} FP64 Jacobi-PCG iterations of the same {n}-DOF system "
                      f"(oracle restatement, OpenMP); oracle assembly took {t_asm:.1f} s"}


def main():
    args = parse()
    if args.profile:
        args.steps, args.warmup, args.no_cpu, args.spmv_reps = 1, 0, True, 5
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU "
              "sharding is described in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)
    from vibestran_amd import HipMixedPcgSolverBackend, _lib
    from vibestran_amd import model as M
    from vibestran_amd.assembly import assemble_model, assembly_stats, dof_map_device

    lib = _lib.load()
    t0 = time.perf_counter()
    m = build_workload(args.config)
    be = HipMixedPcgSolverBackend.try_create(tolerance=args.tol, max_iters=100000)
    if be is None:
        raise RuntimeError("no HIP device")
    eq, n = dof_map_device(be, m)
    K = assemble_model(be, m, eq, n)
    ast = assembly_stats(be)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    log(f"[bench] {m.meta['name']}: n={n} nnz_lower={K.nnz} setup {time.perf_counter()-t0:.1f}s "
        f"(pattern {ast.pattern_ms:.1f} ms, elements {ast.element_ms:.1f} ms)")
    h = be._h
    dF = C.c_void_p()
    du = C.c_void_p()
    nbytes = n * 8
    assert lib.vbh_device_alloc(h, nbytes, C.byref(dF)) == 0
    assert lib.vbh_device_alloc(h, nbytes, C.byref(du)) == 0
    assert lib.vbh_memcpy_h2d(h, dF, F.ctypes.data_as(C.c_void_p), nbytes) == 0
    it = C.c_int(0)
    rel = C.c_double(0.0)

    def solve(prec):
        lib.vbh_set_precision(h, prec)
        rc = lib.vbh_solve_device(h, dF, du, C.byref(it), C.byref(rel))
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        st = _lib.vbh_solve_stats()
        lib.vbh_last_stats(h, C.byref(st))
        return it.value, rel.value, st

    # warmup + timed mixed-precision solves
    for _ in range(args.warmup):
        solve(_lib.VBH_PREC_MIXED)
    walls, devs = [], []
    for k in range(args.steps):
        t = time.perf_counter()
        iters, relres, st = solve(_lib.VBH_PREC_MIXED)
        walls.append(time.perf_counter() - t)
        devs.append(st.solve_ms / 1e3)
        log(f"[bench] mixed solve {k}: {iters} iters, {st.refinements} refreshes, "
            f"true rel {relres:.3e}, {walls[-1]*1e3:.1f} ms wall, {st.solve_ms:.1f} ms device")
    u = np.empty(n)
    lib.vbh_memcpy_d2h(h, u.ctypes.data_as(C.c_void_p), du, nbytes)
    fmt = st.format
    # independent FP64 check of the true residual with the library's FP64 operator
    Ku = be.spmv(u)
    true_rel = float(np.linalg.norm(F - Ku) / np.linalg.norm(F))
    t_mixed = float(np.median(walls))
    iters_mixed = iters

    fp64 = None
    if not args.no_fp64:
        solve(_lib.VBH_PREC_FP64)
        t = time.perf_counter()
        it64, rel64, st64 = solve(_lib.VBH_PREC_FP64)
        t64 = time.perf_counter() - t
        fp64 = {"time_s": t64, "iters": it64, "iters_per_s": it64 / t64, "rel_res": rel64}
        log(f"[bench] fp64 solve: {it64} iters, rel {rel64:.3e}, {t64*1e3:.1f} ms")

    # dominant kernel: the fused FP32 PCG SpMV (search direction + K p + p.Ap),
    # timed with hipEvents on the context stream; plus the update kernel
    def kt(which, reps=args.spmv_reps):
        ms = C.c_double(0.0)
        if lib.vbh_bench_spmv(h, which, reps, C.byref(ms)) != 0:
            raise RuntimeError(_lib.last_error())
        return ms.value
    lib.vbh_memcpy_h2d(h, dF, F.ctypes.data_as(C.c_void_p), nbytes)
    ms32, ms64 = kt(3), kt(2)
    ms_upd32, ms_upd64 = kt(5), kt(4)
    ms_plain32 = kt(1)
    ms_copy = kt(6, 5)
    copy_gbs = 2 * (256 << 20) / (ms_copy * 1e-3) / 1e9
    nnz_full = 2 * K.nnz - n
    # algorithmic bytes per launch, full-CSR-equivalent matrix (value + int32
    # column per nonzero, row pointers) plus the kernel's vector streams:
    #   cg_spmv  reads r, diag, p_old and writes p_new, Ap   (5 streams)
    #   cg_update reads p, Ap, r, diag, u(f64), writes u(f64), r
    spmv32_bytes = nnz_full * (4 + 4) + 4 * (n + 1) + 5 * 4 * n
    spmv64_bytes = nnz_full * (8 + 4) + 4 * (n + 1) + 5 * 8 * n
    upd32_bytes = (4 * 4 + 8 + 8 + 4) * n
    ach32 = spmv32_bytes / (ms32 * 1e-3) / 1e9
    ach64 = spmv64_bytes / (ms64 * 1e-3) / 1e9
    # whole-iteration algorithmic bytes (13 vector streams, FP32 inner)
    b_iter32 = nnz_full * (4 + 4) + 4 * (n + 1) + 13 * n * 4
    iter_gbs = b_iter32 * iters_mixed / (devs[-1]) / 1e9
    log(f"[bench] cg_spmv fp32 {ms32*1e3:.1f} us, fp64 {ms64*1e3:.1f} us; cg_update fp32 "
        f"{ms_upd32*1e3:.1f} us, fp64 {ms_upd64*1e3:.1f} us; plain spmv fp32 {ms_plain32*1e3:.1f} us; "
        f"iteration {devs[-1]/iters_mixed*1e6:.1f} us")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(m, eq, n, F, args.cpu_seconds)
            log(f"[bench] cpu baseline: {cpu['value']:.1f} iters/s on {cpu['cores']} threads")
        except Exception as e:  # the oracle is optional on the box
            log(f"[bench] cpu baseline skipped: {e}")

    value = iters_mixed / t_mixed
    out = {
        "metric": "time-to-1e-8 residual + PCG iters/s, 620k-DOF hex; % HBM roofline vs FP64",
        "value": round(value, 2),
        "unit": "PCG iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 SpMV/Jacobi + f64 residual/solution",
        "data": "synthetic (seeded jittered CHEXA8 block, lognormal E, PLOAD4)",
        "config": {"workload": m.meta["name"], "n_dof": n, "nnz_lower": int(K.nnz),
                   "nnz_full": int(nnz_full), "format": f"BSR{fmt}" if fmt else "CSR",
                   "tol": args.tol, "parallelism": "single GPU"},
        "time_to_1e8_s": round(t_mixed, 4),
        "iters": iters_mixed,
        "true_rel_res": true_rel,
        "fp64": fp64,
        "speedup_vs_fp64": round(fp64["time_s"] / t_mixed, 3) if fp64 else None,
        "roofline": {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach32 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "k_cg_spmv_bsr<float,3> (fused FP32 PCG SpMV)",
                     "bytes_per_launch": int(spmv32_bytes), "us_per_launch": round(ms32 * 1e3, 2)},
        "kernels_us": {"cg_spmv_f32": round(ms32 * 1e3, 2), "cg_spmv_f64": round(ms64 * 1e3, 2),
                       "cg_update_f32": round(ms_upd32 * 1e3, 2),
                       "cg_update_f64": round(ms_upd64 * 1e3, 2),
                       "spmv_plain_f32": round(ms_plain32 * 1e3, 2),
                       "cg_update_f32_GBs": round(upd32_bytes / (ms_upd32 * 1e-3) / 1e9, 1),
                       "cg_spmv_f64_GBs": round(ach64, 1),
                       "stream_copy_256MiB_us": round(ms_copy * 1e3, 2),
                       "stream_copy_GBs": round(copy_gbs, 1)},
        "iteration_roofline": {"bytes_per_iter": int(b_iter32), "achieved": round(iter_gbs, 1),
                               "frac": round(iter_gbs / HBM_PEAK_GBS, 4),
                               "us_per_iter": round(devs[-1] / iters_mixed * 1e6, 2)},
        "assembly_ms": {"pattern": ast.pattern_ms, "elements": ast.element_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
