#!/usr/bin/env python3
"""Benchmark of the linear-static hot path on one MI355X.

Workload (default): CHEXA8 120x40x40 block, +-10% interior jitter, lognormal
E (median 200e3, sigma 0.3), nu=0.3, clamped x=0, PLOAD4 p=1.0 on the top
faces -- 203,401 nodes, 605,160 DOF. K is assembled on the GPU (EAS elements)
and stays in HBM; F is uploaded once. One step = one Jacobi-PCG solve from
u=0 to a true FP64 relative residual ||F-Ku||/||F|| < 1e-8, in FP32 SpMV /
Jacobi with FP64 refreshes and FP64 solution update. The plain FP64 GPU PCG
is timed on the same device matrix for comparison.

Other workloads (--workload tube | simp) print their own JSON line: the
32-load-case block PCG on the CQUAD4 tube, and the adjoint gradients of the
SIMP block. Prints one JSON line per run.
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
    ap.add_argument("--workload", default="hex", choices=["hex", "tube", "simp", "mixed", "deck", "mpc"],
                    help="hex: C1/C2 single-RHS mixed vs FP64 (the headline line); "
                         "tube: 32-load-case block PCG vs sequential solves; "
                         "simp: adjoint compliance / tip-displacement gradients; "
                         "mixed: 1.2M-DOF hex/penta/tet mesh, 8 GRAV+FORCE cases, block "
                         "PCG and compliance adjoint; "
                         "deck: BDF input throughput of the 605k-DOF hex deck (native "
                         "reader vs the Python reader), host only")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solve")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the compact 32-load-case tube leg of the default run")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--profile", action="store_true",
                    help="short run for profilers: one solve per precision, few kernel reps")
    ap.add_argument("--profile-graph", choices=["mixed", "fp64"], default=None,
                    help="short run for profilers of the graph-replayed solve: three solves of "
                         "one precision only (one captured iteration graph in the process), "
                         "then the copy that calibrates the byte counters")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(cfg):
    from vibestran_amd import decks
    if cfg == "c1":
        return decks.config1_hex_cantilever()
    return decks.config2_hex_jitter()


def cpu_threads():
    """Host threads for the CPU legs: the cores this process may run on
    (affinity), capped by OMP_NUM_THREADS (16 on a one-GPU share of the GPU
    host). Returns (threads used, affinity count)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(aff, cap) if cap > 0 else aff), aff


def cpu_leg(model, eq, n, F, u_checks, tol, repeats=3):
    """The oracle on the host: assembles its own FP64 K of the workload, uses it
    to check the GPU solutions (true relative residual ||F - K u|| / ||F|| with
    the oracle's operator, linear_static.cpp:287-304), and times its FP64
    Jacobi-PCG solve to `tol` (the CPU baseline: one warm-up, then the median
    of `repeats` solves; OpenMP over the host threads)."""
    import oracle as O
    nthreads, aff = cpu_threads()
    t0 = time.perf_counter()
    K = O.assemble_solids(n, model.elem_type, model.elem_ptr, model.conn, model.xyz, eq, model.E,
                          model.nu)
    t_asm = time.perf_counter() - t0
    checks = {name: float(O.true_relative_residual(K, u, F)) for name, u in u_checks.items()}
    O.pcg_jacobi(K, F, tol=tol, max_iters=100000, nthreads=nthreads)  # warm-up
    times, iters = [], 0
    for _ in range(repeats):
        t0 = time.perf_counter()
        _, iters, _ = O.pcg_jacobi(K, F, tol=tol, max_iters=100000, nthreads=nthreads)
        times.append(time.perf_counter() - t0)
    t_med = float(np.median(times))
    cpu = {"value": round(iters / t_med, 2), "unit": "PCG iters/s", "cores": nthreads,
           "kind": "port", "time_to_tol_s": round(t_med, 3), "iters": int(iters),
           "times_s": [round(t, 3) for t in times], "host_affinity_cores": aff,
           "sample": f"oracle FP64 Jacobi-PCG of the same {n}-DOF system to {tol:g} (CPU "
                     f"restatement, OpenMP on {nthreads} of {aff} host cores): 1 warm-up + median "
                     f"of {repeats} solves; oracle assembly {t_asm:.1f} s (not in the rate)"}
    return cpu, checks


def profiled_traffic(profile, kernel):
    """From a committed rocprof summary: HBM bytes per launch of `kernel`
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy calibration), its average
    duration in us over the solve (kernel-trace statistics) and its mean
    duration in the back-to-back launches the bench times with events."""
    path = os.path.join(ROOT, "profiles", profile)
    traffic = avg_us = b2b_us = None
    try:
        table = None
        for line in open(path):
            if line.startswith("#"):
                table = None  # a new section: its tables are not read
            elif line.startswith("| kernel | calls | avg us"):
                table = "dur"
            elif line.startswith("| kernel | FETCH_SIZE"):
                table = "traffic"
            elif line.startswith("| launches | mean duration us"):
                table = "b2b"
            elif table == "b2b" and line.startswith("| ") and b2b_us is None:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if cells[0].isdigit():
                    b2b_us = float(cells[1])
            if line.startswith("| `" + kernel + "` |"):
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if table == "traffic":
                    traffic = int((float(cells[3]) + float(cells[4])) * 1e6)
                elif table == "dur" and avg_us is None:
                    avg_us = float(cells[2])
    except OSError:
        return None, None, None, None
    return traffic, avg_us, b2b_us, f"profiles/{profile}"


def main():
    args = parse()
    if args.workload == "tube":
        return bench_tube(args)
    if args.workload == "simp":
        return bench_simp(args)
    if args.workload == "mixed":
        return bench_mixed(args)
    if args.workload == "mpc":
        return bench_mpc(args)
    if args.workload == "deck":
        return bench_deck(args)
    if args.profile:
        args.steps, args.warmup, args.no_cpu, args.spmv_reps = 1, 0, True, 20
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharded solve is a design, not part of this build", file=sys.stderr)
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

    if args.profile_graph:
        prec = _lib.VBH_PREC_MIXED if args.profile_graph == "mixed" else _lib.VBH_PREC_FP64
        for k in range(3):
            iters, relres, st = solve(prec)
            log(f"[bench] {args.profile_graph} solve {k}: {iters} iters, {st.refinements} "
                f"refreshes, rel {relres:.3e}, {st.solve_ms:.1f} ms device")
        ms = C.c_double(0.0)
        lib.vbh_bench_spmv(h, 6, 5, C.byref(ms))
        print(json.dumps({"profile": "hex-graph", "precision": args.profile_graph,
                          "iters": iters, "solve_ms": st.solve_ms}), flush=True)
        return

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
    t_mixed = float(np.median(walls))
    iters_mixed = iters
    refreshes = st.refinements

    fp64 = None
    u64 = None
    if not args.no_fp64:
        solve(_lib.VBH_PREC_FP64)
        t = time.perf_counter()
        it64, rel64, st64 = solve(_lib.VBH_PREC_FP64)
        t64 = time.perf_counter() - t
        u64 = np.empty(n)
        lib.vbh_memcpy_d2h(h, u64.ctypes.data_as(C.c_void_p), du, nbytes)
        fp64 = {"time_s": round(t64, 5), "iters": it64, "iters_per_s": round(it64 / t64, 2),
                "rel_res": rel64}
        log(f"[bench] fp64 solve: {it64} iters, rel {rel64:.3e}, {t64*1e3:.1f} ms")

    # Dominant kernel: the fused FP32 PCG SpMV (K p + the p.Ap, z.Ap and
    # Ap.diag^-1.Ap partials). Its duration is measured live with dispatch
    # start/stop events (hipExtLaunchKernelGGL on the context stream, inside
    # directly launched PCG iterations, the kernel's own begin and end as a
    # profiler reports them); back-to-back launches between one event pair are
    # reported beside.
    def kt(which, reps=args.spmv_reps):
        ms = C.c_double(0.0)
        if lib.vbh_bench_spmv(h, which, reps, C.byref(ms)) != 0:
            raise RuntimeError(_lib.last_error())
        return ms.value

    def op_bytes(prec):
        b = C.c_longlong(0)
        if lib.vbh_operator_stream_bytes(h, prec, C.byref(b)) != 0:
            raise RuntimeError(_lib.last_error())
        return b.value
    lib.vbh_memcpy_h2d(h, dF, F.ctypes.data_as(C.c_void_p), nbytes)
    ms_b2b32, ms_b2b64 = kt(14), kt(13)  # back to back, one event pair around all launches
    ms32, ms64 = kt(8), kt(7)  # inside iterations, the dispatch's own start/stop
    ms_vec32, ms_vec64 = kt(10), kt(9)
    ms_plain32 = kt(1)
    ms_copy = kt(6, 5)
    copy_gbs = 2 * (256 << 20) / (ms_copy * 1e-3) / 1e9
    nnz_full = 2 * K.nnz - n
    n_pad = ((n + fmt - 1) // fmt) * fmt if fmt else n
    # Algorithmic bytes of one fused SpMV launch in the layout it streams: the
    # operator (padded block values, block columns, slice offsets and, for the
    # FP32 stream, the FP64 diagonal node blocks) plus p gathered once, Ap
    # written, and z and diag read for the z.Ap and Ap.diag^-1.Ap partials.
    # The full-CSR equivalent (value + int32 column per nonzero, the same four
    # vectors) is reported beside as an effective figure.
    spmv32_bytes = op_bytes(_lib.VBH_PREC_MIXED) + 4 * 4 * n_pad
    spmv64_bytes = op_bytes(_lib.VBH_PREC_FP64) + 4 * 8 * n_pad
    csr32_bytes = nnz_full * (4 + 4) + 4 * (n + 1) + 4 * 4 * n
    # vector kernel: u (FP64) read and written, p, Ap, diag, r read, r, z, p written
    vec32_bytes = (8 + 4 * 4 + 8 + 3 * 4) * n
    # the live duration: `reps` launches back to back between one event pair on
    # the context stream (the kernel trace of the same loop gives the same
    # per-dispatch duration); the dispatch-event figure inside iterations adds
    # the events' own cost and is reported beside
    ach32 = spmv32_bytes / (ms_b2b32 * 1e-3) / 1e9
    ach64 = spmv64_bytes / (ms_b2b64 * 1e-3) / 1e9
    # whole iteration: the SpMV's bytes plus the vector kernel's
    b_iter32 = spmv32_bytes + vec32_bytes
    iter_gbs = b_iter32 * iters_mixed / (devs[-1]) / 1e9
    log(f"[bench] cg_spmv dispatch events (us): fp32 {ms32*1e3:.2f} fp64 {ms64*1e3:.2f}; back to "
        f"back fp32 {ms_b2b32*1e3:.2f} fp64 {ms_b2b64*1e3:.2f}; cg_vector fp32 {ms_vec32*1e3:.2f} "
        f"fp64 {ms_vec64*1e3:.2f}; "
        f"iteration {devs[-1]/iters_mixed*1e6:.2f} us; fp32 SpMV stream {spmv32_bytes/1e6:.1f} MB")

    # result recovery of the converged u: centroid + corner stresses on the
    # device (device-resident u and outputs), node displacements, CSV files
    post = None
    if not args.profile:
        lib.vbh_memcpy_h2d(h, du, u.ctypes.data_as(C.c_void_p), nbytes)
        from vibestran_amd import results as R
        dcen, dcor, ddisp = C.c_void_p(), C.c_void_p(), C.c_void_p()
        assert lib.vbh_device_alloc(h, m.ne * 80, C.byref(dcen)) == 0
        assert lib.vbh_device_alloc(h, m.ne * 640, C.byref(dcor)) == 0
        assert lib.vbh_device_alloc(h, m.nn * 48, C.byref(ddisp)) == 0
        lib.vbh_recover_stress_device(h, du, dcen, dcor)
        t = time.perf_counter()
        rc = lib.vbh_recover_stress_device(h, du, dcen, dcor)
        t_rec = time.perf_counter() - t
        rc |= lib.vbh_recover_displacements_device(h, du, ddisp)
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        cen = np.empty((m.ne, 10))
        disp = np.empty((m.nn, 6))
        lib.vbh_memcpy_d2h(h, cen.ctypes.data_as(C.c_void_p), dcen, cen.nbytes)
        lib.vbh_memcpy_d2h(h, disp.ctypes.data_as(C.c_void_p), ddisp, disp.nbytes)
        for d in (dcen, dcor, ddisp):
            lib.vbh_device_free(h, d)
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            t = time.perf_counter()
            R.write_csv(os.path.join(td, "c2"), m, [(1, disp, cen)])
            t_csv = time.perf_counter() - t
        post = {"recover_stress_ms": round(t_rec * 1e3, 3), "write_csv_ms": round(t_csv * 1e3, 1),
                "elements": m.ne, "nodes": m.nn}
        log(f"[bench] recovery (centroid + corners) {t_rec*1e3:.2f} ms, CSV {t_csv*1e3:.0f} ms")
    lib.vbh_device_free(h, dF)
    lib.vbh_device_free(h, du)

    # the library's own FP64 operator (always available) and, in the CPU leg,
    # the oracle's independently assembled K
    true_rel_lib = float(np.linalg.norm(F - be.spmv(u)) / np.linalg.norm(F))
    err_vs_fp64 = (float(np.abs(u - u64).max() / np.abs(u64).max()) if u64 is not None else None)
    cpu, checks = None, {}
    if not args.no_cpu:
        try:
            cpu, checks = cpu_leg(m, eq, n, F, {"mixed": u, **({"fp64": u64} if u64 is not None
                                                               else {})}, args.tol)
            log(f"[bench] cpu baseline: {cpu['value']:.1f} iters/s, {cpu['time_to_tol_s']:.2f} s "
                f"to {args.tol:g} on {cpu['cores']} threads; oracle true residuals {checks}")
        except Exception as e:  # the oracle is optional on the box
            log(f"[bench] cpu baseline skipped: {e}")
    be.close()

    c3 = None
    if not args.no_c3 and not args.profile and args.config == "c2":
        c3 = c3_compact(args)

    traffic, prof_us, prof_b2b_us, traffic_src = profiled_traffic(
        "pcg_c2_hex605k.md", f"k_cg_spmv_bsr<float, {fmt}, double>")
    if args.config != "c2":
        traffic, prof_us, prof_b2b_us, traffic_src = None, None, None, None
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
        "time_to_1e8_s": round(t_mixed, 5),
        "iters": iters_mixed,
        "refreshes": refreshes,
        "true_rel_res": checks.get("mixed", true_rel_lib),
        "true_rel_res_operator": "oracle K (CPU FP64 assembly)" if "mixed" in checks
                                 else "library FP64 operator",
        "true_rel_res_library_operator": true_rel_lib,
        "max_norm_err_vs_fp64": err_vs_fp64,
        "fp64": fp64 and {**fp64, "oracle_true_rel_res": checks.get("fp64")},
        "speedup_vs_fp64": round(fp64["time_s"] / t_mixed, 3) if fp64 else None,
        "roofline": {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach32 / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": f"k_cg_spmv_bsr<float, {fmt}, double> (FP32 PCG SpMV + p.Ap, z.Ap, Ap.diag^-1.Ap partials)",
                     "bytes_per_launch": int(spmv32_bytes),
                     "bytes_definition": "sliced-ELL FP32 off-diagonal blocks + FP64 diagonal "
                                         "blocks + block columns + slice offsets + p + Ap "
                                         "+ z + diag",
                     "us_per_launch": round(ms_b2b32 * 1e3, 3),
                     "us_per_launch_dispatch_events": round(ms32 * 1e3, 3),
                     "rocprof_us_per_launch_back_to_back": prof_b2b_us,
                     "rocprof_us_per_launch_in_solve": prof_us,
                     "traffic_source": traffic_src,
                     "frac_of_measured_copy": round(ach32 / copy_gbs, 4),
                     "effective_csr_equivalent": {
                         "bytes_per_launch": int(csr32_bytes),
                         "GBs": round(csr32_bytes / (ms_b2b32 * 1e-3) / 1e9, 1)}},
        "kernels_us": {"cg_spmv_f32": round(ms32 * 1e3, 2), "cg_spmv_f64": round(ms64 * 1e3, 2),
                       "cg_vector_f32": round(ms_vec32 * 1e3, 2),
                       "cg_vector_f64": round(ms_vec64 * 1e3, 2),
                       "cg_spmv_f32_back_to_back": round(ms_b2b32 * 1e3, 2),
                       "cg_spmv_f64_back_to_back": round(ms_b2b64 * 1e3, 2),
                       "spmv_plain_f32": round(ms_plain32 * 1e3, 2),
                       "cg_vector_f32_GBs": round(vec32_bytes / (ms_vec32 * 1e-3) / 1e9, 1),
                       "cg_spmv_f64_bytes": int(spmv64_bytes),
                       "cg_spmv_f64_GBs": round(ach64, 1),
                       "stream_copy_256MiB_us": round(ms_copy * 1e3, 2),
                       "stream_copy_GBs": round(copy_gbs, 1)},
        "iteration_roofline": {"bytes_per_iter": int(b_iter32), "achieved": round(iter_gbs, 1),
                               "frac": round(iter_gbs / HBM_PEAK_GBS, 4),
                               "us_per_iter": round(devs[-1] / iters_mixed * 1e6, 2)},
        "assembly_ms": {"pattern": ast.pattern_ms, "elements": ast.element_ms},
        "results": post,
        "cpu_baseline": cpu,
        "c3_block": c3,
    }
    print(json.dumps(out), flush=True)



def _device_columns(lib, h, A):
    """Column-major host array -> device buffer (caller frees)."""
    ptr = C.c_void_p()
    A = np.asfortranarray(A, dtype=np.float64)
    assert lib.vbh_device_alloc(h, A.nbytes, C.byref(ptr)) == 0
    assert lib.vbh_memcpy_h2d(h, ptr, A.ctypes.data_as(C.c_void_p), A.nbytes) == 0
    return ptr


def c3_compact(args, sample=(0, 8, 16, 24)):
    """The batched workload inside the default run: the CQUAD4 tube (1.92M DOF,
    32 end force/moment load cases) in mixed precision. One block-PCG solve of
    all 32 cases (timed after one warm-up solve, which tunes the SpMM shapes
    and captures the iteration graphs) against single-case mixed solves of the
    sampled cases, scaled to 32. Every case's true relative residual with the
    library's FP64 operator, and the max-norm error of the sampled cases
    against plain FP64 GPU solves of the same cases."""
    from vibestran_amd import HipPcgSolverBackend, _lib, decks
    from vibestran_amd import model as M
    from vibestran_amd.assembly import assemble_model, dof_map_device

    lib = _lib.load()
    t0 = time.perf_counter()
    m = decks.cquad4_tube(n_cases=32)
    k = len(m.loads)
    be = HipPcgSolverBackend.try_create(tolerance=args.tol, max_iters=200000)
    h = be._h
    lib.vbh_set_precision(h, _lib.VBH_PREC_MIXED)
    eq, n = dof_map_device(be, m)
    K = assemble_model(be, m, eq, n)
    F = np.stack([M.assemble_rhs(m, eq, n, lc) for lc in m.loads], axis=1)
    dF, dU = _device_columns(lib, h, F), _device_columns(lib, h, np.zeros_like(F))
    log(f"[bench] c3 tube: n={n}, {k} load cases, setup {time.perf_counter() - t0:.1f} s")
    be.solve_block_device(dF.value, dU.value, k)  # warm-up: tunes the SpMM, builds the graphs
    t = time.perf_counter()
    its, _ = be.solve_block_device(dF.value, dU.value, k)
    t_block = time.perf_counter() - t
    cfg = be.block_config()
    refreshes = be.stats().refinements
    U = np.empty((n, k), order="F")
    lib.vbh_memcpy_d2h(h, U.ctypes.data_as(C.c_void_p), dU, U.nbytes)
    lib.vbh_device_free(h, dF)
    lib.vbh_device_free(h, dU)
    log(f"[bench] c3 block mixed: {t_block:.2f} s, iterations {its.min()}-{its.max()}")
    true = [float(np.linalg.norm(F[:, j] - be.spmv(U[:, j])) / np.linalg.norm(F[:, j]))
            for j in range(k)]
    t_seq, it_seq = 0.0, []
    for j in sample:
        t = time.perf_counter()
        be.solve(K, F[:, j])
        t_seq += time.perf_counter() - t
        it_seq.append(be.last_iteration_count())
    t_seq32 = t_seq * k / len(sample)
    log(f"[bench] c3 sequential mixed sample {list(sample)}: {t_seq:.2f} s -> {t_seq32:.2f} s for {k}")
    lib.vbh_set_precision(h, _lib.VBH_PREC_FP64)
    errs, it64 = [], []
    for j in sample:
        u64 = be.solve(K, F[:, j])
        it64.append(be.last_iteration_count())
        errs.append(float(np.abs(U[:, j] - u64).max() / np.abs(u64).max()))
    be.close()
    return {"workload": m.meta["name"], "n_dof": int(n), "load_cases": k,
            "precision": "f32 SpMM/Jacobi + f64 refresh/solution",
            "block_s": round(t_block, 3), "sequential_32_s": round(t_seq32, 3),
            "speedup": round(t_seq32 / t_block, 3),
            "sequential_sample": list(sample), "sequential_sample_iters": it_seq,
            "block_iters_of_sample": [int(its[j]) for j in sample],
            "block_iters_min_max": [int(its.min()), int(its.max())], "refreshes": refreshes,
            "true_rel_res_max": max(true), "true_rel_res_operator": "library FP64 operator",
            "max_norm_err_vs_fp64_of_sample": errs, "fp64_sample_iters": it64,
            "spmm": {"cols_per_lane": cfg.cols_per_lane, "lanes_per_group": cfg.lanes_per_group,
                     "row_blocks": cfg.row_blocks, "value_path": cfg.value_path,
                     "ms": round(cfg.spmm_ms, 4)}}


def bench_tube(args):
    """C3: CQUAD4 tube R=50, L=400, 400 x 800 elements (1.92M DOF), 32 load
    cases; one block-PCG launch vs single-load-case solves of the same
    operator (a sample of the sequential solves, scaled to 32)."""
    from vibestran_amd import HipPcgSolverBackend, _lib, decks
    from vibestran_amd import model as M
    from vibestran_amd.assembly import assemble_model, dof_map_device

    lib = _lib.load()
    m = decks.cquad4_tube(n_cases=32)
    k = len(m.loads)
    res = {}
    if args.profile:
        return profile_tube(args, m)
    for prec, label in ((_lib.VBH_PREC_FP64, "fp64"), (_lib.VBH_PREC_MIXED, "mixed")):
        be = HipPcgSolverBackend.try_create(tolerance=args.tol, max_iters=200000)
        lib.vbh_set_precision(be._h, prec)
        eq, n = dof_map_device(be, m)
        K = assemble_model(be, m, eq, n)
        F = np.stack([M.assemble_rhs(m, eq, n, lc) for lc in m.loads], axis=1)
        h = be._h
        dF, dU = _device_columns(lib, h, F), _device_columns(lib, h, np.zeros_like(F))
        be.solve_block_device(dF.value, dU.value, k)  # tunes the SpMM, builds the graph
        t = time.perf_counter()
        its, rels = be.solve_block_device(dF.value, dU.value, k)
        t_block = time.perf_counter() - t
        cfg = be.block_config()
        st = be.stats()
        U = np.empty((n, k), order="F")
        lib.vbh_memcpy_d2h(h, U.ctypes.data_as(C.c_void_p), dU, U.nbytes)
        true = [float(np.linalg.norm(F[:, j] - be.spmv(U[:, j])) / np.linalg.norm(F[:, j]))
                for j in (0, k - 1)]
        sample = list(range(0, k, 8))
        t_seq, it_seq = 0.0, []
        for j in sample:
            t = time.perf_counter()
            be.solve(K, F[:, j])
            t_seq += time.perf_counter() - t
            it_seq.append(be.last_iteration_count())
        t_seq32 = t_seq * k / len(sample)
        nnz_full = 2 * K.nnz - n
        sv = 8 if prec == _lib.VBH_PREC_FP64 else 4
        spmm_bytes = nnz_full * (sv + 4) + 4 * (n + 1) + 2 * cfg.ldk * n * sv
        res[label] = {"block_s": round(t_block, 3), "sequential_32_s": round(t_seq32, 3),
                      "speedup": round(t_seq32 / t_block, 3),
                      "block_iters_min_max": [int(its.min()), int(its.max())],
                      "sequential_sample_iters": it_seq,
                      "block_iters_of_sample": [int(its[j]) for j in sample],
                      "refreshes": st.refinements, "true_rel_res_first_last": true,
                      "spmm": {"cols_per_lane": cfg.cols_per_lane,
                               "lanes_per_group": cfg.lanes_per_group,
                               "row_blocks": cfg.row_blocks, "xcd_rows": cfg.xcd_rows,
                               "value_path": cfg.value_path, "ms": round(cfg.spmm_ms, 4),
                               "algorithmic_GBs": round(spmm_bytes / (cfg.spmm_ms * 1e-3) / 1e9, 1)},
                      "ms_per_iteration": round(t_block / int(its.max()) * 1e3, 4)}
        log(f"[bench] tube {label}: block {t_block:.2f} s, 32 sequential ~{t_seq32:.2f} s "
            f"({t_seq32 / t_block:.2f}x), iters {its.min()}-{its.max()}")
        lib.vbh_device_free(h, dF)
        lib.vbh_device_free(h, dU)
        be.close()
    best = max(res.values(), key=lambda r: r["speedup"])
    out = {"metric": "32-RHS block-PCG throughput vs 32 sequential solves, CQUAD4 tube 1.92M DOF",
           "value": best["speedup"], "unit": "x (sequential time / block time)", "n_gpus": 1,
           "steps": 1, "warmup": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 (fp64 mode); f32 SpMM/Jacobi + f64 refresh (mixed mode)",
           "data": "synthetic (seeded tube, 32 random end force/moment cases)",
           "config": {"workload": m.meta["name"], "n_dof": int(n), "load_cases": k, "tol": args.tol},
           "fp64": res["fp64"], "mixed": res["mixed"]}
    print(json.dumps(out), flush=True)


def profile_tube(args, m):
    """Short tube run for profilers: 200 block iterations per precision (the
    solve stops at the iteration limit), then a 256 MiB copy that calibrates
    the HBM byte counters."""
    from vibestran_amd import HipPcgSolverBackend, SolverError, _lib
    from vibestran_amd import model as M
    from vibestran_amd.assembly import assemble_model, dof_map_device

    lib = _lib.load()
    k = len(m.loads)
    for prec in (_lib.VBH_PREC_FP64, _lib.VBH_PREC_MIXED):
        be = HipPcgSolverBackend.try_create(tolerance=args.tol, max_iters=200)
        lib.vbh_set_precision(be._h, prec)
        eq, n = dof_map_device(be, m)
        K = assemble_model(be, m, eq, n)
        F = np.stack([M.assemble_rhs(m, eq, n, lc) for lc in m.loads], axis=1)
        try:
            be.solve_block(K, F)
        except SolverError:
            pass  # stopped at the 200-iteration limit, as intended
        cfg = be.block_config()
        log(f"[bench] tube profile prec={prec}: {be.last_block_iterations.max()} iterations, "
            f"SpMM cw={cfg.cols_per_lane} lg={cfg.lanes_per_group} values={cfg.value_path} "
            f"{cfg.spmm_ms*1e3:.1f} us")
        ms = C.c_double(0.0)
        lib.vbh_bench_spmv(be._h, 6, 5, C.byref(ms))
        be.close()
    print(json.dumps({"profile": "tube", "load_cases": k}), flush=True)


def bench_simp(args):
    """C4: CHEXA8 80x40x20 SIMP block; compliance and tip-displacement
    gradients w.r.t. every element density from one 2-column block solve
    ([F, e_tip]) and two adjoint contraction launches; the GPU gradients are
    compared with the oracle's FP64 CPU adjoint on all elements."""
    import oracle as O
    from vibestran_amd import HipPcgSolverBackend, _lib, decks
    from vibestran_amd import model as M
    from vibestran_amd.assembly import assemble_model, dof_map_device, element_sensitivity

    m = decks.config4_simp()
    be = HipPcgSolverBackend.try_create(tolerance=1e-10, max_iters=400000)
    eq, n = dof_map_device(be, m)
    tip = int(eq[m.meta["tip_node"], 2])
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    R = np.zeros((n, 2))
    R[:, 0] = F
    R[tip, 1] = 1.0
    dE = 3.0 * m.meta["density"] ** 2 * (m.meta["E0"] - m.meta["Emin"])
    times = []
    for step in range(2):
        t0 = time.perf_counter()
        K = assemble_model(be, m, eq, n)
        t1 = time.perf_counter()
        U = be.solve_block(K, R)
        t2 = time.perf_counter()
        gC = element_sensitivity(be, m, 0, U[:, 0], U[:, 0]) * dE
        gT = element_sensitivity(be, m, 0, U[:, 0], U[:, 1]) * dE
        t3 = time.perf_counter()
        times.append((t1 - t0, t2 - t1, t3 - t2))
    its = be.last_block_iterations
    mE = M.Model(**{**m.__dict__})
    oC = O.element_sensitivity(n, mE, eq, 0, U[:, 0], U[:, 0]) * dE
    oT = O.element_sensitivity(n, mE, eq, 0, U[:, 0], U[:, 1]) * dE
    errC = float(np.abs(gC - oC).max() / np.abs(oC).max())
    errT = float(np.abs(gT - oT).max() / np.abs(oT).max())
    ta, ts, tg = times[-1]
    out = {"metric": "adjoint dC/drho and du_tip/drho for all elements, CHEXA8 80x40x20 SIMP",
           "value": round(ta + ts + tg, 4), "unit": "s (assembly + 2-column solve + 2 contractions)",
           "n_gpus": 1, "steps": 1, "warmup": 1, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded SIMP densities)",
           "config": {"workload": m.meta["name"], "n_dof": int(n), "elements": int(m.ne), "tol": 1e-10},
           "assembly_s": round(ta, 4), "solve_s": round(ts, 4), "contractions_s": round(tg, 4),
           "iterations": [int(x) for x in its],
           "max_rel_diff_vs_cpu_adjoint": {"compliance": errC, "tip": errT},
           "compliance": float(F @ U[:, 0]), "u_tip": float(U[tip, 0])}
    print(json.dumps(out), flush=True)


def bench_deck(args):
    """Input throughput: the 605k-DOF jittered hex workload written as a BDF
    deck (one MAT1 + PSOLID per element, SPC1 THRU, 4,800 PLOAD4; free
    field, and a mixed large/small/free field copy), read by the native
    reader (vbh_deck_read) and, for the free-field copy, by the pure-Python
    reader; both must give the same model. Host only."""
    import tempfile
    from vibestran_amd import bdf, deck, decks

    m = decks.config2_hex_jitter()
    res = {}
    with tempfile.TemporaryDirectory() as td:
        for fmt in ("free", "mixed"):
            path = os.path.join(td, f"c2_{fmt}.bdf")
            text = decks.bdf_text(m, fmt)
            with open(path, "w") as fh:
                fh.write(text)
            size = os.path.getsize(path)
            nlines = text.count("\n")
            del text
            deck.read_deck(path)  # warm the page cache
            t = time.perf_counter()
            a = deck.read_deck(path)
            t_native = time.perf_counter() - t
            r = {"bytes": size, "lines": nlines, "native_s": round(t_native, 4),
                 "native_MBs": round(size / t_native / 1e6, 1)}
            if fmt == "free":
                t = time.perf_counter()
                b = bdf.read_bdf(path)
                t_py = time.perf_counter() - t
                same = (np.array_equal(a.xyz, b.xyz) and np.array_equal(a.conn, b.conn)
                        and np.array_equal(a.E, b.E) and np.array_equal(a.loads[0].pload4.p,
                                                                        b.loads[0].pload4.p))
                r.update({"python_s": round(t_py, 3), "speedup": round(t_py / t_native, 1),
                          "same_model": bool(same)})
                base = r
            res[fmt] = r
            log(f"[bench] deck {fmt}: {size/1e6:.1f} MB, {nlines} lines, native {t_native:.3f} s"
                + (f", python {r['python_s']:.2f} s" if "python_s" in r else ""))
    out = {"metric": "BDF input throughput, 605k-DOF hex deck (192k CHEXA + 192k MAT1/PSOLID)",
           "value": base["native_MBs"], "unit": "MB/s (native reader, free field)", "n_gpus": 0,
           "steps": 1, "warmup": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 text parsing",
           "data": "synthetic (the hex120x40x40 jittered workload written as BDF text)",
           "config": {"workload": m.meta["name"] + "-bdf", "nodes": m.nn, "elements": m.ne},
           "free": res["free"], "mixed": res["mixed"]}
    print(json.dumps(out), flush=True)


def bench_mpc(args):
    """MPC elimination at the 605k-DOF hex size: the end face tied to its
    centre grid (5,040 equations) and the y = 0 face on inclined slides
    (4,879 equations). Times the numbering + vbh_set_mpc (host), the device
    assembly with K_red = T^T K T, and the FP64 and mixed solves of the
    reduced system; the reduced residual T^T (K u - F) is checked through the
    device element forces."""
    import scipy.sparse as sp
    from vibestran_amd import HipMixedPcgSolverBackend, HipPcgSolverBackend, decks
    from vibestran_amd import results as R
    from vibestran_amd.assembly import (analysis_numbering, assemble_model, assemble_rhs_device,
                                        assembly_stats, mpc_expand, mpc_transform)

    m = decks.hex_block_mpc(120, 40, 40)
    res = {}
    for name, cls in (("fp64", HipPcgSolverBackend), ("mixed", HipMixedPcgSolverBackend)):
        be = cls.try_create(tolerance=1e-8, max_iters=400000)
        t0 = time.perf_counter()
        eq, n, n_red = analysis_numbering(be, m)
        t_num = time.perf_counter() - t0
        t0 = time.perf_counter()
        K = assemble_model(be, m, eq, n)
        t_asm = time.perf_counter() - t0
        ast = assembly_stats(be)
        F = assemble_rhs_device(be, m, eq, n_red, m.loads[0])
        be.solve(K, F)  # warm-up (graph capture)
        t0 = time.perf_counter()
        u = be.solve(K, F)
        t_solve = time.perf_counter() - t0
        st = be.stats()
        f = R.element_forces(be, m, u)
        mask = eq >= 0
        Ku = np.zeros(n)
        Ff = np.zeros(n)
        np.add.at(Ku, eq[mask], f[mask])
        np.add.at(Ff, eq[mask], m.loads[0].nodal[mask])
        tp, tc, tv = mpc_transform(be)
        r = sp.csr_matrix((tv, tc, tp), shape=(n, n_red)).T @ (Ku - Ff)
        res[name] = {"solve_s": round(t_solve, 4), "iterations": int(st.iterations),
                     "format": int(st.format), "reduced_rel_residual": float(np.linalg.norm(r) / np.linalg.norm(F)),
                     "numbering_and_set_mpc_s": round(t_num, 3), "assembly_s": round(t_asm, 3),
                     "mpc_transform_ms": round(ast.mpc_ms, 2), "nnz_reduced": int(ast.nnz)}
        log(f"[bench] mpc {name}: n {n} -> {n_red}, transform {ast.mpc_ms:.1f} ms, solve "
            f"{t_solve*1e3:.1f} ms in {st.iterations} iterations (format {st.format})")
        be.close()
    out = {"metric": "linear-static solve with 9.9k MPC equations eliminated, 605k-DOF hex",
           "value": res["mixed"]["solve_s"], "unit": "s (mixed-precision solve of K_red to 1e-8)",
           "n_gpus": 1, "steps": 1, "warmup": 1, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 transform; f32 SpMV + f64 refresh solve",
           "data": "synthetic (hex120x40x40 jittered, tie + slide MPC equations)",
           "config": {"workload": m.meta["name"], "n_full": int(n), "n_reduced": int(n_red),
                      "mpc_equations": int(m.mpc.n)},
           "fp64": res["fp64"], "mixed": res["mixed"]}
    print(json.dumps(out), flush=True)


def bench_mixed(args):
    """C5: ~1.2M-DOF mixed mesh (CHEXA8 / CPENTA6 / CTETRA4, two materials),
    8 load cases of GRAV plus nodal forces assembled on the GPU, one block
    PCG (FP32 + FP64 refresh) over the 8 cases vs single-case solves (two
    of them, scaled to 8), and the compliance adjoint w.r.t. every element E
    for case 0, compared with the oracle's CPU adjoint."""
    import oracle as O
    from vibestran_amd import HipMixedPcgSolverBackend, _lib, decks
    from vibestran_amd.assembly import (assemble_model, assemble_rhs_device, assembly_stats,
                                        dof_map_device, element_sensitivity)

    lib = _lib.load()
    t0 = time.perf_counter()
    m = decks.mixed_block()
    t_deck = time.perf_counter() - t0
    be = HipMixedPcgSolverBackend.try_create(tolerance=args.tol, max_iters=400000)
    h = be._h
    t0 = time.perf_counter()
    eq, n = dof_map_device(be, m)
    K = assemble_model(be, m, eq, n)
    t_asm = time.perf_counter() - t0
    ast = assembly_stats(be)
    k = len(m.loads)
    dF, dU = C.c_void_p(), C.c_void_p()
    assert lib.vbh_device_alloc(h, k * n * 8, C.byref(dF)) == 0
    assert lib.vbh_device_alloc(h, k * n * 8, C.byref(dU)) == 0
    t0 = time.perf_counter()
    for j, lc in enumerate(m.loads):
        assemble_rhs_device(be, m, eq, n, lc, F_device=dF.value + j * n * 8)
    t_rhs = time.perf_counter() - t0
    F = np.empty((n, k), order="F")
    lib.vbh_memcpy_d2h(h, F.ctypes.data_as(C.c_void_p), dF, F.nbytes)
    be.solve_block_device(dF.value, dU.value, k)  # tunes and captures the graph
    t0 = time.perf_counter()
    its, rels = be.solve_block_device(dF.value, dU.value, k)
    t_block = time.perf_counter() - t0
    cfg = be.block_config()
    U = np.empty((n, k), order="F")
    lib.vbh_memcpy_d2h(h, U.ctypes.data_as(C.c_void_p), dU, U.nbytes)
    true = [float(np.linalg.norm(F[:, j] - be.spmv(U[:, j])) / np.linalg.norm(F[:, j]))
            for j in range(k)]
    t_seq, it_seq = 0.0, []
    for j in (0, k // 2):
        t = time.perf_counter()
        be.solve(K, F[:, j])
        t_seq += time.perf_counter() - t
        it_seq.append(be.last_iteration_count())
    t_seq8 = t_seq * k / 2
    t0 = time.perf_counter()
    g = element_sensitivity(be, m, 0, U[:, 0], U[:, 0])
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    go = O.element_sensitivity(n, m, eq, 0, U[:, 0], U[:, 0])
    t_cpu_adj = time.perf_counter() - t0
    err = float(np.abs(g - go).max() / np.abs(go).max())
    lib.vbh_device_free(h, dF)
    lib.vbh_device_free(h, dU)
    log(f"[bench] mixed mesh n={n}: block {t_block:.2f} s vs ~{t_seq8:.2f} s sequential, "
        f"iters {its.min()}-{its.max()}, adjoint {t_adj*1e3:.1f} ms (CPU {t_cpu_adj:.1f} s, diff {err:.1e})")
    out = {"metric": "8-RHS block-PCG to 1e-8 + compliance adjoint, ~1.2M-DOF hex/penta/tet mesh",
           "value": round(t_block, 3), "unit": "s (block solve of 8 load cases)", "n_gpus": 1,
           "steps": 1, "warmup": 1, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 SpMM/Jacobi + f64 refresh/solution",
           "data": "synthetic (seeded jittered mixed mesh, GRAV + random nodal forces)",
           "config": {"workload": m.meta["name"], "n_dof": int(n), "elements": m.meta["counts"],
                      "load_cases": k, "tol": args.tol, "nnz_lower": int(K.nnz)},
           "deck_s": round(t_deck, 3), "dof_map_and_assembly_s": round(t_asm, 3),
           "assembly_ms": {"pattern": ast.pattern_ms, "elements": ast.element_ms},
           "rhs_8_cases_s": round(t_rhs, 4),
           "block_iters_min_max": [int(its.min()), int(its.max())],
           "sequential_sample_iters": it_seq, "sequential_8_s": round(t_seq8, 3),
           "speedup_vs_sequential": round(t_seq8 / t_block, 3),
           "true_rel_res_max": max(true),
           "spmm": {"cols_per_lane": cfg.cols_per_lane, "lanes_per_group": cfg.lanes_per_group,
                    "row_blocks": cfg.row_blocks, "xcd_rows": cfg.xcd_rows,
                    "value_path": cfg.value_path, "ms": round(cfg.spmm_ms, 4)},
           "adjoint_dC_dE_s": round(t_adj, 4), "cpu_adjoint_s": round(t_cpu_adj, 2),
           "adjoint_max_rel_diff_vs_cpu": err}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
