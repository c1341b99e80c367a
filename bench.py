#!/usr/bin/env python3
"""Benchmark of the aggregation-AMG preconditioned CG solve on one MI355X.

Workload: 3D 7-point steady heat conduction on a 128^3 cell grid
(2,097,152 unknowns, 14,581,760 nonzeros), log-normal cell conductivity
(mu 0, sigma 1, seed 1), harmonic face averages, Dirichlet T=0 / T=1 on the
x faces.  One step = AMG setup + PCG solve (V(2,2) Jacobi, omega 2/3,
coarsest <= 256 rows by dense LU) to a 1e-8 FP64 relative residual, with the
matrix and right-hand side already resident in HBM.

  chosen: FP32 V-cycle (levels stored in FP32) under FP64 CG, FP64 coarse LU
  plain:  the same method with the cycle in FP64

value = seconds per chosen setup+solve (lower is better); the plain time and
the ratio are reported beside it.  roofline = the level-0 FP64 SpMV kernel
the solve runs (sliced ELL with 16-bit column offsets; 216,924,164
algorithmic CSR bytes per launch at 128^3) timed with HIP events on the
library stream; the CSR kernel is timed beside it.  cpu_baseline = the CPU oracle (FP64, OpenMP) on the
same problem (FP64 setup + PCG solve), timed on the host cores.
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "setup+solve s to 1e-8 FP64 rel. residual at 128³, chosen vs plain; SpMV GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--cpu-nx", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-graph", action="store_true", help="launch each iteration's kernels without hipGraph")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the level-0 SpMV launches and a copy kernel of known size (for rocprofv3 --pmc)")
    ap.add_argument("--workload", choices=["heat3d", "convdiff", "fem", "batch", "learned", "infer"], default="heat3d",
                    help="convdiff: 96^3 upwind convection-diffusion, BiCGStab with V/W/F cycles, FP64 and FP32 "
                         "levels, cell Peclet 10/30/100; fem: P1 tets on 64^3 cells, 64 right-hand sides in one "
                         "block CG vs 64 single solves, and the k=64 SpMM; batch: 32 independent heat problems "
                         "(edges U{16..96}) in one segmented solve vs 32 separate solves; learned: held-out "
                         "rotated anisotropic and jump-coefficient Q1 problems, strength vs learned affinities "
                         "(each one JSON line of its own); infer: the committed network's inference "
                         "(tf_gnn_infer: graph build, two message-passing layers, scorer) on the 128^3 log-normal heat "
                         "operator and the held-out rotated anisotropic Q1 operator")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the solver runs on a single "
              "MI355X and multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from thermofem_amd import problems as P
    from thermofem_amd._lib import check, lib
    from thermofem_amd.device import Context, DeviceArray, DeviceCsr
    from thermofem_amd.solver import AmgOptions, Cg, KrylovOptions, SolveStatus

    if args.workload == "convdiff":
        convdiff(args)
        return
    if args.workload == "fem":
        fem_block(args)
        return
    if args.workload == "batch":
        batch(args)
        return
    if args.workload == "learned":
        learned(args)
        return
    if args.workload == "infer":
        infer(args)
        return

    nx = args.nx
    log(f"generating heat3d log-normal {nx}^3")
    (rp, ci, v), b = P.heat3d_lognormal(nx, seed=1)
    n, nnz = len(rp) - 1, len(ci)
    A = DeviceCsr(rp, ci, v)
    db = DeviceArray.from_host(b)
    dx = DeviceArray(n, np.float64)
    # the initial guess x0 = 0, resident in HBM like the other inputs
    dzero = DeviceArray.from_host(np.zeros(n))

    def make(precision):
        amg = AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2.0 / 3.0,
                         strength_threshold=0.25, matching_rounds=8, level_precision=precision)
        return Cg(KrylovOptions(relative_tolerance=1e-8, use_graph=not args.no_graph), amg)

    def step(solver):
        dx.copy_from(dzero)
        solver.setup_device(n, nnz, *A.args())
        return solver.apply_device(db, dx)

    if args.spmv_only:
        spmv_only(args, n, nnz, A, v, np, P, check, lib, Context, DeviceArray)
        return

    results = {}
    for name, prec in (("chosen", 32), ("plain", 64)):
        solver = make(prec)
        # The first step of each configuration is reported on its own: it
        # pays the tuner's timings for the level shapes, the memory pool's
        # first allocations, graph capture and, for the first configuration,
        # the kernels' code-object loading.
        cold = None
        for w in range(args.warmup):
            log(f"{name}: warmup {w}")
            t0 = time.perf_counter()
            step(solver)
            if cold is None:
                cold = (time.perf_counter() - t0, *solver.timings())
        times, setups, solves = [], [], []
        res = None
        for k in range(args.steps):
            t0 = time.perf_counter()
            res = step(solver)
            t1 = time.perf_counter()
            s_setup, s_apply = solver.timings()
            if cold is None:
                cold = (t1 - t0, s_setup, s_apply)
            times.append(t1 - t0)
            setups.append(s_setup)
            solves.append(s_apply)
            log(f"{name}: step {k} {1e3 * (t1 - t0):.2f} ms (setup {1e3 * s_setup:.2f}, solve {1e3 * s_apply:.2f}) "
                f"it={res.iterations} rel={res.relative_residual:.3e} {res.status.name}")
            if res.status != SolveStatus.Converged:
                raise SystemExit(f"{name} solve did not converge: {res}")
        view = solver.amg()
        for d, (r_, z_, h_) in enumerate(view.level_statistics()):
            log(f"{name}: level {d} rows {r_} nnz {z_} ({z_ / r_:.1f}/row) aggregates by size {h_}")
        results[name] = dict(total=sum(times), mean=statistics.fmean(times), median=statistics.median(times),
                             setup=statistics.median(setups), solve=statistics.median(solves), res=res,
                             x=dx.get(), levels=view.level_count(), stats=view.level_statistics(), cold=cold)
        del solver

    ch, pl = results["chosen"], results["plain"]
    err = float(np.linalg.norm(ch["x"] - pl["x"]) / np.linalg.norm(pl["x"]))

    # Level-0 SpMV on the library stream, HIP events around the launches.
    ctx = Context(0)
    y64 = DeviceArray(n, np.float64)
    y32 = DeviceArray(n, np.float32)
    xv = DeviceArray.from_host(P.uniform(11, n))
    x32 = DeviceArray.from_host(P.uniform(11, n).astype(np.float32))
    v32 = DeviceArray.from_host(v.astype(np.float32))
    sell, bits, stored = make_sell(ctx, n, nnz, A, check, lib)
    stencil, terms, arrays = make_stencil(ctx, n, nnz, A, check, lib)
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    check(lib().tf_event_create(C.byref(ev0)))
    check(lib().tf_event_create(C.byref(ev1)))

    # Every launch is timed cold: a 512 MB read (the library's asum, plain
    # loads) first replaces the level-0 data in the L2s and the 256 MB
    # Infinity Cache with clean lines, so the rate is the HBM-bound one.  Back
    # to back, the ~180 MB the kernel reads would stay in the Infinity Cache
    # (the "warm" figure below); a copy as the flush would leave dirty lines
    # whose write-back the timed kernel would pay.
    flush_buf = DeviceArray(1 << 26, np.float64)
    L = lib()

    def time_launches(launch, reps, cold=True):
        ms = C.c_float()
        sink = C.c_double()
        total = 0.0
        for r in range(reps + 2):
            if cold:
                check(L.tf_asum_f64(ctx.h, 1 << 26, flush_buf.ptr, C.byref(sink)))
            check(L.tf_event_record(ctx.h, ev0))
            check(launch())
            check(L.tf_event_record(ctx.h, ev1))
            check(L.tf_event_elapsed_ms(ev0, ev1, C.byref(ms)))
            if r >= 2:
                total += ms.value
        return total / reps * 1e-3

    reps = args.spmv_reps
    st64 = st32 = st32w = None
    if stencil:
        st64 = time_launches(lambda: L.tf_stencil_spmv_f64(ctx.h, stencil, xv.ptr, y64.ptr), reps)
        st32 = time_launches(lambda: L.tf_stencil_spmv_f32(ctx.h, stencil, x32.ptr, y32.ptr), reps)
        st32w = time_launches(lambda: L.tf_stencil_spmv_f32(ctx.h, stencil, x32.ptr, y32.ptr), reps, cold=False)
        check(L.tf_stencil_destroy(stencil))
    t64 = time_launches(lambda: L.tf_sell_spmv_f64(ctx.h, sell, xv.ptr, y64.ptr), reps)
    t32 = time_launches(lambda: L.tf_sell_spmv_f32(ctx.h, sell, x32.ptr, y32.ptr), reps)
    c64 = time_launches(lambda: L.tf_spmv_csr_f64(ctx.h, n, nnz, A.rp.ptr, A.ci.ptr, A.v.ptr, xv.ptr, y64.ptr), reps)
    c32 = time_launches(lambda: L.tf_spmv_csr_f32(ctx.h, n, nnz, A.rp.ptr, A.ci.ptr, v32.ptr, x32.ptr, y32.ptr),
                        reps)
    warm32 = time_launches(lambda: L.tf_sell_spmv_f32(ctx.h, sell, x32.ptr, y32.ptr), reps, cold=False)
    del flush_buf
    check(L.tf_sell_destroy(sell))
    bytes64 = P.spmv_bytes(n, nnz, 8)
    bytes32 = P.spmv_bytes(n, nnz, 4)
    gbs64 = bytes64 / t64 / 1e9
    gbs32 = bytes32 / t32 / 1e9
    # bytes the sliced-ELL kernel itself must move: values + 16-bit offsets
    # (padding included), row lengths, slice offsets, x once and y
    moved64 = stored * (8 + bits // 8) + n + 4 * ((n + 63) // 64 + 1) + 16 * n
    log(f"spmv sliced-ELL f64 {t64 * 1e6:.1f} us {gbs64:.0f} GB/s; f32 {t32 * 1e6:.1f} us {gbs32:.0f} GB/s; "
        f"CSR f64 {c64 * 1e6:.1f} us, f32 {c32 * 1e6:.1f} us")
    # The solve's level-0 products run in the structured-stencil storage when
    # the operator has one offset list (it does here): the roofline kernel.
    if st64 is not None:
        moved_st = n * (8 * arrays + 2 + 16)
        variant = 0 if (terms == 7 and arrays == 4) else (1 if terms <= 8 else 2)
        roof_t, roof_kernel = st64, (f"k_stencil<double,0,{variant}> level-0 FP64 SpMV, structured-stencil storage "
                                     f"({terms} offsets, {arrays} value arrays"
                                     f"{', -1/+1 neighbours from adjacent lanes' if variant == 0 else ''})")
        log(f"spmv stencil f64 {st64 * 1e6:.1f} us {bytes64 / st64 / 1e9:.0f} GB/s; f32 {st32 * 1e6:.1f} us")
    else:
        roof_t, roof_kernel = t64, "k_sell<double,short,0> level-0 FP64 SpMV"

    cpu = None
    oracle_err = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        cnx = args.cpu_nx
        log(f"cpu baseline: oracle FP64 AMG-PCG at {cnx}^3")
        (crp, cci, cv), cb = P.heat3d_lognormal(cnx, seed=1)
        t0 = time.perf_counter()
        host = O.Amg(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2.0 / 3.0).setup((crp, cci, cv))
        xo, st, it, _, rel = O.krylov("cg", host, (crp, cci, cv), cb)
        tcpu = time.perf_counter() - t0
        if cnx == nx:
            oracle_err = {"chosen": float(np.linalg.norm(ch["x"] - xo) / np.linalg.norm(xo)),
                          "plain": float(np.linalg.norm(pl["x"] - xo) / np.linalg.norm(xo))}
        sample = (f"the full {nx}^3 workload" if cnx == nx else
                  f"a {cnx}^3 instance ({(nx // cnx) ** 3}x smaller than the {nx}^3 workload)")
        cpu = {"value": round(tcpu, 4), "unit": "s", "cores": int(O.lib().tfo_threads()), "kind": "port",
               "sample": f"{sample}: CPU oracle FP64 AMG setup + PCG solve to 1e-8 "
                         f"(status {st}, {it} iterations, rel {rel:.2e})"}
        log(f"cpu baseline {tcpu:.2f} s, {it} iterations")

    out = {
        "metric": METRIC,
        "value": round(ch["median"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ch["total"] / args.steps * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": f"heat3d log-normal {nx}^3 AMG-PCG setup+solve to 1e-8 (FP32 V(2,2) cycle, FP64 CG)",
            "n": n, "nnz": nnz, "levels": ch["levels"], "coarse_rows": 256,
        },
        "plain_fp64_s": round(pl["median"], 6),
        "speedup_vs_plain": round(pl["median"] / ch["median"], 4),
        "chosen": {"setup_s": round(ch["setup"], 6), "solve_s": round(ch["solve"], 6),
                   "iterations": ch["res"].iterations, "relative_residual": ch["res"].relative_residual},
        "plain": {"setup_s": round(pl["setup"], 6), "solve_s": round(pl["solve"], 6),
                  "iterations": pl["res"].iterations, "relative_residual": pl["res"].relative_residual},
        "cold_first_step": {k: {"step_s": round(r["cold"][0], 6), "setup_s": round(r["cold"][1], 6),
                                "solve_s": round(r["cold"][2], 6)} for k, r in (("chosen", ch), ("plain", pl))},
        "cold_note": "first step of each configuration in a fresh process (chosen runs first): tuner timings, "
                     "first pool allocations, graph capture, and for chosen the kernels' first loads",
        "rel_error_chosen_vs_plain_fp64": err,
        "rel_error_vs_fp64_oracle": oracle_err,
        "spmv": {"f64_us": round(t64 * 1e6, 2), "f64_GBs": round(gbs64, 1), "f32_us": round(t32 * 1e6, 2),
                 "f32_GBs": round(gbs32, 1), "bytes_f64": bytes64, "bytes_f32": bytes32,
                 "format": f"sliced ELL, slices of 64 rows, {bits}-bit columns, {stored} stored entries",
                 "f64_bytes_moved": moved64, "f64_moved_GBs": round(moved64 / t64 / 1e9, 1),
                 "csr_f64_us": round(c64 * 1e6, 2), "csr_f32_us": round(c32 * 1e6, 2),
                 "timing": "each launch after a 512 MB read that evicts L2 and Infinity Cache (HBM-bound rate)",
                 "f32_us_warm": round(warm32 * 1e6, 2),
                 "f32_GBs_warm": round(bytes32 / warm32 / 1e9, 1),
                 "stencil": None if st64 is None else {
                     "f64_us": round(st64 * 1e6, 2), "f64_GBs": round(bytes64 / st64 / 1e9, 1),
                     "f32_us": round(st32 * 1e6, 2), "f32_GBs": round(bytes32 / st32 / 1e9, 1),
                     "f32_us_warm": round(st32w * 1e6, 2), "offsets": terms, "value_arrays": arrays,
                     "f64_bytes_moved": moved_st, "f64_moved_GBs": round(moved_st / st64 / 1e9, 1)}},
        "roofline": {"bound": "hbm", "achieved": round(bytes64 / roof_t / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(bytes64 / roof_t / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic_bytes(nx, st64 is not None), "kernel": roof_kernel,
                     "algorithmic_bytes": bytes64,
                     "own_bytes": moved_st if st64 is not None else moved64,
                     "achieved_own_bytes": round((moved_st if st64 is not None else moved64) / roof_t / 1e9, 1),
                     "frac_own_bytes": round((moved_st if st64 is not None else moved64) / roof_t / 1e9 /
                                             HBM_PEAK_GBS, 4),
                     "note": "achieved = the CSR-algorithmic bytes (nnz*(8+4) + 4(n+1) + 16n) over the cold launch "
                             "time; the stencil storage reads about half of what that formula counts (no column "
                             "indices, one half of the symmetric values), so achieved_own_bytes / frac_own_bytes "
                             "give the rate of the bytes the kernel must move"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def convdiff(args):
    """96^3 convection-diffusion (upwind, 8 Fourier velocity modes) at cell
    Peclet 10, 30 and 100: BiCGStab + AMG (V(1,1) Jacobi omega 2/3, coarsest
    <= 256 rows by dense LU) to a 1e-8 FP64 relative residual with V, W and F
    cycles, levels in FP64 and in FP32.  Median seconds of setup+solve over
    --steps runs after --warmup, one JSON line."""
    import numpy as np
    from thermofem_amd import problems as P
    from thermofem_amd.device import DeviceArray, DeviceCsr
    from thermofem_amd.solver import AmgOptions, BiCgStab, Cycle, KrylovOptions, SolveStatus

    nx = 96 if args.nx == 128 else args.nx
    rows = []
    for pe in (10.0, 30.0, 100.0):
        log(f"generating convection-diffusion {nx}^3 Pe {pe:g}")
        (rp, ci, v), b = P.convection_diffusion3d(nx, pe)
        n, nnz = len(rp) - 1, len(ci)
        A = DeviceCsr(rp, ci, v)
        db = DeviceArray.from_host(b)
        dx = DeviceArray(n, np.float64)
        zeros = np.zeros(n)
        for cyc in (Cycle.V, Cycle.W, Cycle.F):
            for prec in (64, 32):
                s = BiCgStab(KrylovOptions(relative_tolerance=1e-8, use_graph=not args.no_graph),
                             AmgOptions(coarse_rows=256, jacobi_weight=2.0 / 3.0, cycle=cyc, level_precision=prec))
                times, setups, solves, res = [], [], [], None
                for k in range(args.warmup + args.steps):
                    dx.upload(zeros)
                    t0 = time.perf_counter()
                    s.setup_device(n, nnz, *A.args())
                    res = s.apply_device(db, dx)
                    t1 = time.perf_counter()
                    if k >= args.warmup:
                        st, sa = s.timings()
                        times.append(t1 - t0)
                        setups.append(st)
                        solves.append(sa)
                x = dx.get()
                true_rel = float(np.linalg.norm(b - csr_matvec(rp, ci, v, x)) / np.linalg.norm(b))
                row = {"peclet": pe, "cycle": cyc.name, "levels_fp": prec, "s": round(statistics.median(times), 6),
                       "setup_s": round(statistics.median(setups), 6), "solve_s": round(statistics.median(solves), 6),
                       "iterations": res.iterations, "status": res.status.name, "true_rel_residual": true_rel}
                log(json.dumps(row))
                rows.append(row)
                if res.status != SolveStatus.Converged:
                    log(f"warning: Pe {pe} {cyc.name} FP{prec} did not converge")
                del s
    best = min((r for r in rows if r["status"] == "Converged"), key=lambda r: r["s"], default=None)
    print(json.dumps({"metric": f"BiCGStab+AMG setup+solve s to 1e-8 at {nx}^3 convection-diffusion",
                      "unit": "s", "higher_is_better": False, "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "dtype": "f64 / f32 levels", "data": "synthetic",
                      "config": {"workload": f"upwind convection-diffusion {nx}^3, 8 Fourier velocity modes, "
                                             "cell Peclet 10/30/100, V(1,1) omega 2/3, coarsest <= 256",
                                 "n": nx ** 3},
                      "best": best, "runs": rows}), flush=True)


def fem_block(args):
    """P1 tetrahedral heat operator M/dt + K (64^3 cells x 6 Kuhn tets,
    274,625 nodes, 4,018,753 nonzeros, dt 1e-3, log-normal k_tet) with 64
    right-hand sides: AMG (V(2,2) Jacobi omega 2/3, coarsest <= 256 by dense
    LU) + CG to 1e-8 on all 64 at once (row-major n x 64 blocks, SpMM) against
    64 single-vector solves with the same hierarchy; plus the k = 64 FP64
    SpMM alone (algorithmic bytes nnz*12 + 4(n+1) + 2*64*n*8)."""
    import numpy as np
    from thermofem_amd import problems as P
    from thermofem_amd._lib import check, lib
    from thermofem_amd.device import Context, DeviceArray, DeviceCsr
    from thermofem_amd.solver import AmgOptions, Cg, KrylovOptions, SolveStatus

    nc, k = 64, 64
    log(f"generating P1 tet heat operator {nc}^3 cells")
    rp, ci, v = P.fem_p1_heat(nc)
    n, nnz = len(rp) - 1, len(ci)
    A = DeviceCsr(rp, ci, v)
    B = P.block_rhs(n, k)
    dB = DeviceArray.from_host(B.ravel())
    dX = DeviceArray(n * k, np.float64)
    zeros = np.zeros(n * k)
    out = {"metric": f"setup + CG solve s of {k} right-hand sides to 1e-8, P1 tets {nc}^3 cells",
           "unit": "s", "higher_is_better": False, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "data": "synthetic", "config": {"workload": f"P1 tet M/dt+K {nc}^3 cells x 6 Kuhn tets, {k} RHS",
                                           "n": n, "nnz": nnz, "k": k}}
    for prec in (32, 64):
        s = Cg(KrylovOptions(relative_tolerance=1e-8, use_graph=not args.no_graph),
               AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2.0 / 3.0,
                          level_precision=prec))
        times, setups, solves, res = [], [], [], None
        for it in range(args.warmup + args.steps):
            dX.upload(zeros)
            t0 = time.perf_counter()
            s.setup_device(n, nnz, *A.args())
            t1 = time.perf_counter()
            res = s.apply_block_device(k, dB, dX)
            t2 = time.perf_counter()
            if it >= args.warmup:
                times.append(t2 - t0)
                setups.append(t1 - t0)
                solves.append(t2 - t1)
        X = dX.get().reshape(n, k)
        worst = max(float(np.linalg.norm(B[:, j] - csr_matvec(rp, ci, v, X[:, j])) / np.linalg.norm(B[:, j]))
                    for j in range(k))
        its = [r.iterations for r in res]
        ok = all(r.status == SolveStatus.Converged for r in res)
        block = {"s": round(statistics.median(times), 6), "setup_s": round(statistics.median(setups), 6),
                 "solve_s": round(statistics.median(solves), 6), "iterations_min_max": [min(its), max(its)],
                 "all_converged": ok, "worst_true_rel_residual": worst}
        log(f"fem fp{prec} block: {block}")
        # the same 64 systems one at a time (same solver and hierarchy)
        x1 = DeviceArray(n, np.float64)
        single_solve, single_its = 0.0, []
        for j in range(k):
            b1 = DeviceArray.from_host(np.ascontiguousarray(B[:, j]))
            x1.upload(np.zeros(n))
            t0 = time.perf_counter()
            r1 = s.apply_device(b1, x1)
            single_solve += time.perf_counter() - t0
            single_its.append(r1.iterations)
        single = {"solve_s": round(single_solve, 6), "s": round(statistics.median(setups) + single_solve, 6),
                  "iterations_min_max": [min(single_its), max(single_its)]}
        log(f"fem fp{prec} 64 single solves: {single}")
        out[f"levels_fp{prec}"] = {"block": block, "single_solves": single,
                                   "block_speedup": round(single["s"] / block["s"], 3)}
        del s
    # k = 64 FP64 SpMM with HIP events on the library stream
    ctx = Context(0)
    Y = DeviceArray(n * k, np.float64)
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    check(lib().tf_event_create(C.byref(ev0)))
    check(lib().tf_event_create(C.byref(ev1)))
    L = lib()

    v32 = DeviceArray.from_host(v.astype(np.float32))
    B32 = DeviceArray.from_host(B.ravel().astype(np.float32))
    Y32 = DeviceArray(n * k, np.float32)

    def time_spmm(launch):
        for _ in range(3):
            check(launch())
        reps = args.spmv_reps
        check(L.tf_event_record(ctx.h, ev0))
        for _ in range(reps):
            check(launch())
        check(L.tf_event_record(ctx.h, ev1))
        ms = C.c_float()
        check(L.tf_event_elapsed_ms(ev0, ev1, C.byref(ms)))
        return ms.value / reps * 1e-3

    for name, s_, launch in (
            ("spmm_k64_f64", 8, lambda: L.tf_spmm_csr_f64(ctx.h, n, nnz, A.rp.ptr, A.ci.ptr, A.v.ptr, k, dB.ptr,
                                                          Y.ptr)),
            ("spmm_k64_f32", 4, lambda: L.tf_spmm_csr_f32(ctx.h, n, nnz, A.rp.ptr, A.ci.ptr, v32.ptr, k, B32.ptr,
                                                          Y32.ptr))):
        t = time_spmm(launch)
        nbytes = nnz * (s_ + 4) + 4 * (n + 1) + 2 * k * n * s_
        out[name] = {"us": round(t * 1e6, 2), "algorithmic_bytes": nbytes, "GBs": round(nbytes / t / 1e9, 1),
                     "frac_of_8TBs": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"{name} {t * 1e6:.1f} us {nbytes / t / 1e9:.0f} GB/s")
    out["value"] = out["levels_fp32"]["block"]["s"]
    print(json.dumps(out), flush=True)


def batch(args):
    """32 independent log-normal heat problems, cube edges ~ U{16..96} (seed
    3), as one block-diagonal matrix: one AMG setup + segmented CG (FP32
    V(2,2) cycle, FP64 CG, to 1e-8 per system), the solve as per-kernel
    launches (`value`: captured iterations replayed between control-block
    reads) and as one persistent launch, against 32 separate setup + solve
    runs of the same solver configuration."""
    import numpy as np
    from thermofem_amd import problems as P
    from thermofem_amd.device import DeviceArray, DeviceCsr
    from thermofem_amd.solver import AmgOptions, Cg, KrylovOptions, SolveStatus

    log("generating 32 heat problems")
    (rp, ci, v), b, off, edges = P.heat3d_batch()
    n, nnz = len(rp) - 1, len(ci)
    A = DeviceCsr(rp, ci, v)
    db = DeviceArray.from_host(b)
    dx = DeviceArray(n, np.float64)
    amg = AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=2.0 / 3.0, level_precision=32)
    s = Cg(KrylovOptions(relative_tolerance=1e-8, use_graph=not args.no_graph), amg)

    def worst_of(x):
        worst = 0.0
        for k in range(len(edges)):
            lo, hi = off[k], off[k + 1]
            rs = b[lo:hi] - csr_matvec(rp[lo:hi + 1] - rp[lo], ci[rp[lo]:rp[hi]] - lo, v[rp[lo]:rp[hi]], x[lo:hi])
            worst = max(worst, float(np.linalg.norm(rs) / np.linalg.norm(b[lo:hi])))
        return worst

    # per-kernel launches (captured iterations replayed between control reads)
    # and the whole solve as one persistent launch, each timed as setup + solve
    results = {}
    for mode in ("launched", "persistent"):
        times, solves, res = [], [], None
        for it in range(args.warmup + args.steps):
            dx.upload(np.zeros(n))
            t0 = time.perf_counter()
            s.setup_device(n, nnz, *A.args())
            t1 = time.perf_counter()
            res = s.apply_segmented_device(off, db, dx, persistent=mode == "persistent")
            t2 = time.perf_counter()
            if it >= args.warmup:
                times.append(t2 - t0)
                solves.append(t2 - t1)
        its = [r.iterations for r in res]
        results[mode] = {"s": round(statistics.median(times), 6), "solve_s": round(statistics.median(solves), 6),
                         "iterations_min_max": [min(its), max(its)],
                         "all_converged": all(r.status == SolveStatus.Converged for r in res),
                         "worst_true_rel_residual": worst_of(dx.get())}
        log(f"segmented ({mode}): {results[mode]}")
    seg = results["launched"]
    del s
    # the same systems one at a time (setup + solve each)
    one = Cg(KrylovOptions(relative_tolerance=1e-8, use_graph=not args.no_graph), amg)
    total, sits = 0.0, []
    for k in range(len(edges)):
        lo, hi = int(off[k]), int(off[k + 1])
        Ak = DeviceCsr(rp[lo:hi + 1] - rp[lo], ci[rp[lo]:rp[hi]] - lo, v[rp[lo]:rp[hi]])
        bk = DeviceArray.from_host(np.ascontiguousarray(b[lo:hi]))
        xk = DeviceArray(hi - lo, np.float64)
        for rep in range(2):  # second run timed (first warms the tuner caches)
            xk.upload(np.zeros(hi - lo))
            t0 = time.perf_counter()
            one.setup_device(hi - lo, int(rp[hi] - rp[lo]), *Ak.args())
            r1 = one.apply_device(bk, xk)
            t1 = time.perf_counter()
        total += t1 - t0
        sits.append(r1.iterations)
    single = {"s": round(total, 6), "iterations_min_max": [min(sits), max(sits)]}
    log(f"32 separate solves: {single}")
    print(json.dumps({"metric": "setup + CG solve s of 32 independent heat problems (edges U{16..96}) to 1e-8",
                      "unit": "s", "higher_is_better": False, "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "dtype": "f64+f32", "data": "synthetic",
                      "config": {"workload": "32 log-normal heat3d cubes, edges U{16..96} seed 3, segmented CG + "
                                             "AMG FP32 V(2,2)", "n": n, "nnz": nnz, "edges": edges.tolist()},
                      "value": seg["s"], "segmented": seg, "segmented_persistent": results["persistent"],
                      "separate": single,
                      "speedup": round(single["s"] / seg["s"], 3)}), flush=True)


def learned(args):
    """Held-out Q1 diffusion problems on nc^3 cells (default 128: 127^3
    unknowns, 27-point rows): rotated anisotropic tensors diag(1, eps, eps)
    per 16^3-cell block (seeds 7 and 8) and a four-material jump cube (seed
    9).  AMG-PCG (FP32 V(2,2) cycle, FP64 CG, Jacobi omega 0.45 for the
    anisotropic operators whose D^-1 A reaches ~3.9, 2/3 otherwise) to 1e-8
    with the strength weights (plain), the committed network's weights
    (learned = 1) and the learned hierarchy only where its PCG estimate wins
    (learned = 2): setup and solve seconds, iterations, which hierarchy the
    fallback kept."""
    import numpy as np
    from thermofem_amd import problems as P
    from thermofem_amd.device import DeviceArray, DeviceCsr
    from thermofem_amd.gnn import AffinityModel
    from thermofem_amd.solver import AmgOptions, Cg, KrylovOptions

    nc = args.nx
    model = AffinityModel.trained()
    cases = [("anisotropic seed 7", lambda: P.anisotropic_diffusion(nc, seed=7), 0.45),
             ("anisotropic seed 8", lambda: P.anisotropic_diffusion(nc, seed=8), 0.45),
             ("four-material jump seed 9", lambda: P.jump_diffusion(nc, seed=9), 2.0 / 3.0)]
    rows, fallbacks = [], 0
    for name, make, omega in cases:
        log(f"generating {name} {nc}^3 cells")
        (rp, ci, v), b = make()
        n, nnz = len(rp) - 1, len(ci)
        A = DeviceCsr(rp, ci, v)
        db = DeviceArray.from_host(b)
        dx = DeviceArray(n, np.float64)
        for mode in (0, 1, 2):
            s = Cg(KrylovOptions(relative_tolerance=1e-8, max_iterations=3000),
                   AmgOptions(coarse_rows=256, pre_sweeps=2, post_sweeps=2, jacobi_weight=omega, level_precision=32,
                              learned=mode))
            view = s.amg()
            if mode:
                view.set_affinity_model(model)
            times = []
            for it in range(args.warmup + args.steps):
                dx.upload(np.zeros(n))
                t0 = time.perf_counter()
                s.setup_device(n, nnz, *A.args())
                res = s.apply_device(db, dx)
                t1 = time.perf_counter()
                if it >= args.warmup:
                    times.append((t1 - t0, *s.timings()))
            used, rl, rpl = view.learned_info()
            fallbacks += mode == 2 and not used
            row = {"problem": name, "learned": mode, "s": round(statistics.median(t[0] for t in times), 6),
                   "setup_s": round(statistics.median(t[1] for t in times), 6),
                   "solve_s": round(statistics.median(t[2] for t in times), 6), "iterations": res.iterations,
                   "status": res.status.name, "learned_hierarchy_used": used, "pcg_estimate_learned": rl,
                   "pcg_estimate_plain": rpl}
            log(json.dumps(row))
            rows.append(row)
            del s
    print(json.dumps({"metric": f"AMG-PCG setup+solve s to 1e-8, Q1 diffusion {nc}^3 cells, strength vs learned "
                                "affinities", "unit": "s", "higher_is_better": False, "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "dtype": "f64+f32 (network bf16/fp32)",
                      "data": "synthetic", "config": {"workload": f"held-out Q1 anisotropic/jump {nc}^3 cells",
                                                      "n": (nc - 1) ** 3},
                      "fallback_rate": fallbacks / len(cases), "runs": rows}), flush=True)


def infer(args):
    """Learned-affinity inference alone: tf_gnn_infer on the 128^3 log-normal
    heat operator and on the rotated anisotropic Q1 operator (128^3 cells,
    seed 7), each call timed to completion (the entry point synchronises)."""
    import numpy as np
    from thermofem_amd import problems as P
    from thermofem_amd._lib import check, lib
    from thermofem_amd.device import Context, DeviceArray, DeviceCsr
    from thermofem_amd.gnn import AffinityModel

    model = AffinityModel.trained()
    ctx = Context(0)
    nc = args.nx
    for name, make in (("heat3d log-normal", lambda: P.heat3d_lognormal(nc, seed=1)[0]),
                       ("anisotropic seed 7 (Q1)", lambda: P.anisotropic_diffusion(nc, seed=7)[0])):
        a = make()
        n, nnz = len(a[0]) - 1, len(a[1])
        A = DeviceCsr(*a)
        w = DeviceArray(nnz, np.float64)
        times = []
        for it in range(args.warmup + args.steps):
            ctx.sync()
            t0 = time.perf_counter()
            check(lib().tf_gnn_infer(model.handle, n, nnz, *A.args(), w.ptr))  # returns when done
            if it >= args.warmup:
                times.append(time.perf_counter() - t0)
        print(json.dumps({"metric": f"learned-affinity inference s, {name} {nc}^3", "value":
                          round(statistics.median(times), 6), "unit": "s", "higher_is_better": False, "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "dtype": "f32 (bf16 weights, split-bf16 MFMA)",
                          "data": "synthetic", "config": {"workload": f"{name} {nc}^3", "n": n, "nnz": nnz}}),
              flush=True)


def csr_matvec(rp, ci, v, x):
    """y = A x on the host in FP64 (for the recomputed residual)."""
    import numpy as np
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return np.bincount(rows, weights=v * x[ci], minlength=len(rp) - 1)


def make_sell(ctx, n, nnz, A, check, lib):
    """Sliced-ELL handle of the level-0 matrix (the format the solve uses)."""
    h = C.c_void_p()
    check(lib().tf_sell_create(ctx.h, n, nnz, A.rp.ptr, A.ci.ptr, A.v.ptr, 0.3, C.byref(h)))
    bits, stored = C.c_int(), C.c_int64()
    check(lib().tf_sell_info(h, C.byref(bits), C.byref(stored), None))
    return h, bits.value, stored.value


def make_stencil(ctx, n, nnz, A, check, lib):
    """Structured-stencil handle of the level-0 matrix (None if it has no single offset list)."""
    h = C.c_void_p()
    if lib().tf_stencil_create(ctx.h, n, nnz, A.rp.ptr, A.ci.ptr, A.v.ptr, C.byref(h)) != 0:
        return None, 0, 0
    terms, arrays = C.c_int(), C.c_int()
    check(lib().tf_stencil_info(h, C.byref(terms), C.byref(arrays), None))
    return h, terms.value, arrays.value


def traffic_bytes(nx, stencil=True):
    """HBM bytes per level-0 SpMV launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (copy-kernel calibrated), for the 128^3 workload only."""
    name = "spmv_stencil_f64_level0_heat3d128_pmc.json" if stencil else "spmv_sell_f64_level0_heat3d128_pmc.json"
    path = os.path.join(ROOT, "profiles", name)
    if nx != 128 or not os.path.exists(path):
        return None
    with open(path) as f:
        return int(round(json.load(f)["traffic_bytes"]))


def spmv_only(args, n, nnz, A, v, np, P, check, lib, Context, DeviceArray):
    """Level-0 FP64 SpMV in sliced ELL (k_sell<double,short,0>) and in the
    structured-stencil storage (k_stencil<double,0,0> for the mirrored 7-point
    operator) x reps each, and a plain copy kernel
    (k_convert<double,double>, 2*8*n bytes per launch) x reps, for counter runs."""
    ctx = Context(0)
    xv = DeviceArray.from_host(P.uniform(11, n))
    y = DeviceArray(n, np.float64)
    sell, _, _ = make_sell(ctx, n, nnz, A, check, lib)
    stencil, _, _ = make_stencil(ctx, n, nnz, A, check, lib)
    for _ in range(args.spmv_reps):
        check(lib().tf_sell_spmv_f64(ctx.h, sell, xv.ptr, y.ptr))
    if stencil:
        for _ in range(args.spmv_reps):
            check(lib().tf_stencil_spmv_f64(ctx.h, stencil, xv.ptr, y.ptr))
    big = DeviceArray(8 * n, np.float64)
    big2 = DeviceArray(8 * n, np.float64)
    for _ in range(args.spmv_reps):
        check(lib().tf_copy_f64(ctx.h, 8 * n, big.ptr, big2.ptr))
    ctx.sync()
    print(json.dumps({"spmv_launches": args.spmv_reps, "spmv_bytes": P.spmv_bytes(n, nnz, 8),
                      "copy_launches": args.spmv_reps, "copy_bytes_read": 8 * 8 * n, "copy_bytes_written": 8 * 8 * n}),
          flush=True)


if __name__ == "__main__":
    main()
