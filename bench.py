#!/usr/bin/env python3
"""Throughput of YETI's 3D NURBS solid (U1) FP64 assembly on one MI355X.

Workload (one step): the full drop-in output of sys_linmat_lindef_static
(build_system_elemStorage.f90:24-315) for a 10x1x1 cantilever meshed with 64x64x64 degree-3
NURBS elements (262,144 elements, 902,289 dofs, 4.86e9 COO entries): every element's
upper-triangle stiffness blocks (FP64 values + int32 row/col ids, 77.7 GB) and the traction
right-hand side F, written to HBM. Inputs (control points, connectivity, knots, weights) are
resident in HBM before the timed region starts.

Prints one JSON line on stdout; progress goes to stderr. ``roofline`` describes the dominant
kernel (the element kernel with index emission), each launch timed with HIP events on the
stream it runs on; ``cpu_baseline`` is YETI's own OpenMP Fortran routine (oracle/_ref) on a
bounded sample of the same workload, on this host's cores. Extra objects report the other
parts of the metric on the same GPU: ``serving`` (a Poisson sample of the serving workload:
solves/s and p50/p99 latency, mixed precision vs FP64, vs plain one-at-a-time FP64 solves),
``gradient`` (compliance shape gradient of a 32^3 p=2 patch against central differences) and
``mixed_precision`` (128^3 p=2 matrix-free FP32 solve with FP64 refinement vs FP64, and the
matrix-free apply alone in DOF-applies/s against the HBM roofline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
# The serving section keeps 16 request streams busy; HIP maps streams onto GPU_MAX_HW_QUEUES
# hardware queues (4 by default), so this process asks for 16. Must be set before the HIP
# runtime starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = "assembled elem/s + served solves/s, p99 latency vs plain FP64 at rel err ≤1e-8"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(nel, p, slab, runs, threads):
    """YETI's sys_linmat_lindef_static_omp (compiled reference, oracle/_ref) on this host's cores:
    an nel x nel x slab patch of the benchmark's generator (the same element type, degree,
    quadrature and jitter as the benchmark patch, so the same work per element; a prefix of the
    benchmark patch itself would leave all but one thread idle, as the routine's
    `!$omp parallel do` hands out contiguous element ranges) and, in full, the 16x4x4 p=2
    cantilever. Each figure is the median of ``runs`` calls after one warm-up, output allocation
    included. Falls back to the C restatement (oracle/liboracle.so) when the reference library is
    absent."""
    import numpy as np
    from oracle import oracle as O
    from yeti_amd.geometry import make_box_patch
    kind = "reference" if O.ref_available() else "port"

    def timed(args, tag):
        run = (lambda: O.ref_sys_linmat_omp(args, num_threads=threads)) if kind == "reference" else \
            (lambda: O.oracle_sys_linmat(args, nthreads=threads))
        run()
        ts = []
        for k in range(runs):
            t0 = time.perf_counter()
            run()
            ts.append(time.perf_counter() - t0)
            log("cpu baseline %s: call %d of %d, %.2f s" % (tag, k + 1, runs, ts[-1]))
        return float(np.median(ts)), ts

    bp = make_box_patch(nel=(nel, nel, slab), p=p)
    t_s, ts_s = timed(bp.f2py_args(), "%dx%dx%d p=%d" % (nel, nel, slab, p))
    small = make_box_patch(nel=(16, 4, 4), p=2)
    t_small, _ = timed(small.f2py_args(), "16x4x4 p=2")
    return {"value": bp.nb_elem / t_s, "unit": "elem/s", "cores": threads, "kind": kind,
            "sample": "%dx%dx%d p=%d patch of the benchmark's generator (%d elements), "
                      "sys_linmat_lindef_static_omp incl. output allocation, median of %d calls after a "
                      "warm-up: %.2f s (runs %s)" % (nel, nel, slab, p, bp.nb_elem, runs, t_s,
                                                    ", ".join("%.2f" % t for t in ts_s)),
            "cantilever_16x4x4_p2": {"elements": small.nb_elem, "seconds": t_small, "value": small.nb_elem / t_small,
                             "unit": "elem/s"}}


def read_traffic(path):
    """HBM bytes per launch of the element kernel from a committed rocprofv3 PMC summary
    (line 'traffic_bytes_per_launch: N'), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def bench_serving(n, rate, seed, budget, plain_n):
    """The serving workload: a Poisson stream of n requests (seeded random patches, 4-48
    elements per direction, degree 2-4) at `rate` per second, served by the batched server
    (every active request advanced by shared launches) in mixed precision and in FP64, both to
    1e-8 on the FP64 residual; the plain FP64 GPU path (each request alone, eager launches, one
    after the other) on the first plain_n requests."""
    import numpy as np
    import torch
    from yeti_amd.serving import SolveServer, request_stream
    reqs = request_stream(n=n, rate=rate, seed=seed)
    t0 = time.perf_counter()
    for r in reqs:
        r.model()
    log("serving: %d requests generated in %.1f s (mean %.0f dofs)"
        % (n, time.perf_counter() - t0, np.mean([3 * r.nb_cp for r in reqs])))
    out = {}
    xs = {}
    for prec in ("mixed", "fp64"):
        srv = SolveServer(budget_bytes=budget, precision=prec, chunk=16, rtol=1e-8, batched=True)
        t1 = time.perf_counter()
        rep = srv.run(reqs, keep_solutions=True, realtime=True,
                      progress=lambda d, t: (d % 250 == 0) and log("  serving %s: %d / %d done, %.1f s"
                                                                   % (prec, d, t, time.perf_counter() - t1)))
        xs[prec] = rep.pop("x")
        its = [d["iters"] for d in rep["requests"].values()]
        out[prec] = {"solves_per_s": rep["solves_per_s"], "latency_s": rep["latency_s"],
                     "completed": rep["completed"], "failed": rep["failed"], "at_fp64_floor": rep["at_fp64_floor"],
                     "mean_iterations": float(np.mean(its)), "makespan_s": rep["makespan_s"],
                     "peak_arena_gb": rep["peak_arena_bytes"] / 1e9, "batch_steps": rep["steps"],
                     "host_admission_s": rep["host_admission_s"], "step_wait_s": rep["step_wait_s"]}
        log("serving %s: %.2f solves/s, p50 %.3f s, p99 %.3f s, failed %d, mean iterations %.0f"
            % (prec, rep["solves_per_s"], rep["latency_s"]["p50"], rep["latency_s"]["p99"], rep["failed"],
               np.mean(its)))
        srv.close()
        del srv
        torch.cuda.empty_cache()
    err = max(float(np.linalg.norm(xs["mixed"][k] - xs["fp64"][k]) / np.linalg.norm(xs["fp64"][k])) for k in xs["fp64"])
    # capacity: the first cap_n requests of the stream all admitted at t = 0, served to the end
    # (solves/s at saturation, where the offered rate no longer bounds the throughput)
    cap_n = min(n, 400)
    for prec in ("mixed", "fp64"):
        srv = SolveServer(budget_bytes=budget, precision=prec, chunk=16, rtol=1e-8, batched=True)
        rep = srv.run(reqs[:cap_n], realtime=False)
        out[prec]["capacity"] = {"requests": cap_n, "solves_per_s": rep["solves_per_s"],
                                 "makespan_s": rep["makespan_s"], "failed": rep["failed"]}
        log("serving %s capacity: %d requests at t = 0 in %.2f s, %.2f solves/s"
            % (prec, cap_n, rep["makespan_s"], rep["solves_per_s"]))
        srv.close()
        del srv
        torch.cuda.empty_cache()
    # plain FP64 GPU path: each request alone, one after the other, eager launches, one stream
    sub = reqs[:plain_n]
    lat = []
    t0 = time.perf_counter()
    for r in sub:
        srv = SolveServer(budget_bytes=budget // 4, precision="fp64", nstreams=1, chunk=16, rtol=1e-8, use_graph=False)
        ts = time.perf_counter()
        srv.run([r], realtime=False)
        lat.append(time.perf_counter() - ts)
        srv.close()
        del srv
    dt = time.perf_counter() - t0
    out["plain_fp64"] = {"solves_per_s": len(sub) / dt, "requests": len(sub),
                         "latency_s": {"p50": float(np.median(lat)), "p99": float(np.percentile(lat, 99)),
                                       "max": float(max(lat))}}
    log("plain fp64 (one at a time, %d requests): %.2f solves/s" % (len(sub), len(sub) / dt))
    out["max_rel_err_mixed_vs_fp64"] = err
    out["speedup_mixed_vs_plain_fp64"] = out["mixed"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]
    out["speedup_mixed_vs_served_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["capacity_mixed_vs_served_fp64"] = (out["mixed"]["capacity"]["solves_per_s"] /
                                            out["fp64"]["capacity"]["solves_per_s"])
    out["capacity_mixed_vs_plain_fp64"] = out["mixed"]["capacity"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]
    out["workload"] = ("%d requests, Poisson arrivals at %.1f/s, elements per direction log-uniform in [4,48], "
                       "p in {2,3,4}, E~U[1e3,1e5], nu~U[0.2,0.4], traction~U[0.1,10], solved to 1e-8 "
                       "(FP64 residual), fast-diagonalisation PCG, arena %.0f GB; served: one batch of sessions "
                       "advanced by shared launches (graph per 16 iterations); plain FP64: the first %d requests "
                       "solved alone, one after the other, eager launches; capacity: the first %d requests "
                       "admitted at t = 0" % (n, rate, budget / 1e9, len(sub), cap_n))
    return out


def bench_solve_64cube(nel, p, jacobi_iters):
    """The solve of the assembled patch (64^3 p=3 by default): device CSR of K + K^T, its nodal-block SpMV against the
    HBM roofline, Jacobi PCG to 1e-10 (2-norm) with the true residual and the FP64 rounding floor
    eps ||(|A||x|)|| / ||b|| of the final iterate, and the fast-diagonalisation PCG session (FP64,
    matrix-free) to 1e-10, its solution checked with the assembled CSR."""
    import numpy as np
    import torch
    from yeti_amd.assembly import Assembler
    from yeti_amd.geometry import make_box_patch
    from yeti_amd.serving import SolveServer, Request
    from yeti_amd.solver import StiffnessCSR, free_mask_from_args
    bp = make_box_patch(nel=(nel, nel, nel), p=p)
    asm = Assembler(bp.args)
    K = torch.empty(asm.nb_data, dtype=torch.float64, device="cuda")
    asm.run_values(K)
    F = asm.run_rhs()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    csr = StiffnessCSR(asm)
    torch.cuda.synchronize()
    t_pat = time.perf_counter() - t0
    vals = csr.values(K)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        csr.values(K, out=vals)
    torch.cuda.synchronize()
    t_val = (time.perf_counter() - t0) / 3
    del K
    torch.cuda.empty_cache()
    x = torch.randn(csr.nrows, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    csr.spmv(vals, x, y)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(10):
        csr.spmv(vals, x, y)
    e1.record(stream)
    torch.cuda.synchronize()
    spmv_ms = e0.elapsed_time(e1) / 10
    ncp = csr.nrows // 3
    nblocks = csr.nnz // 9
    # values, one block column per 3x3 block, the block row pointers, x read, y written
    spmv_bytes = 8 * csr.nnz + 4 * nblocks + 8 * (ncp + 1) + 16 * csr.nrows
    log("%d^3 p=%d: CSR pattern %.2f s (nnz %d), values %.1f ms, SpMV %.3f ms (%.0f GB/s)"
        % (nel, p, t_pat, csr.nnz, t_val * 1e3, spmv_ms, spmv_bytes / (spmv_ms * 1e-3) / 1e9))
    mask = free_mask_from_args(bp.args)
    maskd = mask.to(torch.float64)
    bnorm = float((F * maskd).norm())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u, it, rel = csr.pcg(vals, F, mask, rtol=1e-10, maxit=jacobi_iters, check_every=100, raise_on_maxit=False)
    torch.cuda.synchronize()
    t_jac = time.perf_counter() - t0
    true_rel = float(((F - csr.spmv(vals, u)) * maskd).norm()) / bnorm
    floor = float(np.finfo(np.float64).eps) * float((csr.spmv_abs(vals, u) * maskd).norm()) / bnorm
    log("assembled patch, Jacobi PCG: %d iterations in %.2f s (%.3f ms/iteration), relative residual %.3e (true %.3e, "
        "FP64 floor eps||A||x|| / ||b|| %.3e)" % (it, t_jac, t_jac / max(it, 1) * 1e3, rel, true_rel, floor))
    req = Request(0, 0.0, (nel, nel, nel), p, 210000.0, 0.3, -1.0, 20261202)
    req.args = bp.args
    srv = SolveServer(budget_bytes=64 << 30, precision="fp64", nstreams=1, chunk=32, rtol=1e-10, max_iters=20000)
    t0 = time.perf_counter()
    rep = srv.run([req], keep_solutions=True, realtime=False)
    t_fdm = time.perf_counter() - t0
    d = rep["requests"][0]
    uf = torch.as_tensor(rep["x"][0], device="cuda")
    srv.close()
    rel_csr = float(((F - csr.spmv(vals, uf)) * maskd).norm()) / bnorm
    du = float((uf - u).norm() / u.norm())
    log("assembled patch, fast-diagonalisation PCG (FP64, matrix-free): %d iterations in %.2f s, residual %.3e "
        "(assembled CSR: %.3e), |u_fdm - u_jacobi| / |u_jacobi| %.2e" % (d["iters"], t_fdm, d["relres"], rel_csr, du))
    out = {"workload": "%d^3 p=%d cantilever (%d dofs): device CSR of K + K^T from the FP64 COO values, "
                       "Jacobi PCG and fast-diagonalisation PCG to 1e-10" % (nel, p, csr.nrows),
           "csr_pattern_s": t_pat, "csr_values_ms": t_val * 1e3, "nnz": csr.nnz,
           "spmv": {"ms": spmv_ms, "kernel": "block_spmv_kernel (nodal 3x3 blocks, one wave per control point)",
                    "algorithmic_bytes": spmv_bytes,
                    "roofline": {"bound": "hbm", "achieved": spmv_bytes / (spmv_ms * 1e-3) / 1e9,
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": spmv_bytes / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}},
           "jacobi_pcg": {"iterations": it, "max_iterations": jacobi_iters, "seconds": t_jac,
                          "ms_per_iteration": t_jac / max(it, 1) * 1e3, "relres": rel, "true_relres": true_rel,
                          "fp64_floor_eps_absA_absx_over_b": floor, "converged": rel <= 1e-10},
           "fdm_pcg_fp64": {"iterations": d["iters"], "seconds": t_fdm, "relres": d["relres"],
                            "relres_assembled_csr": rel_csr, "stopped_at_fp64_floor": d["floor"],
                            "failed": d["failed"], "rel_diff_vs_jacobi_solution": du}}
    del csr, vals, asm
    torch.cuda.empty_cache()
    return out


def bench_gradient(nfd, seed=7):
    import numpy as np
    import torch
    from yeti_amd.geometry import make_box_patch
    from yeti_amd.gradient import CompliancePatch, compliance
    bp = make_box_patch(nel=(32, 32, 32), p=2)
    patch = CompliancePatch(bp.args, solver="csr", rtol=1e-12)
    X = patch.coords0.clone().cuda()
    Xg = X.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = compliance(Xg, patch)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g = Xg.grad.detach().cpu().numpy()
    rng = np.random.default_rng(seed)
    idx = rng.choice(X.numel(), nfd, replace=False)
    h = 1e-6
    errs = []
    for i in idx:
        Xp = X.clone(); Xp[i] += h
        Xm = X.clone(); Xm[i] -= h
        fd = (compliance(Xp, patch).item() - compliance(Xm, patch).item()) / (2 * h)
        errs.append(abs(fd - g[i]) / max(abs(g).max(), 1e-300))
        if len(errs) % 10 == 0:
            log("gradient: %d / %d central differences done" % (len(errs), nfd))
    compliance(X, patch)
    log("gradient: forward %.3f s, backward %.3f s, max |fd - g| / max|g| %.2e over %d coordinates"
        % (t1 - t0, t2 - t1, max(errs), nfd))
    return {"workload": "32x32x32 p=2 cantilever (117,912 dofs), J = F.u, FP64 assembly + CSR Jacobi PCG to 1e-12",
            "forward_s": t1 - t0, "backward_s": t2 - t1, "J": J.item(), "fd_step": h, "fd_coordinates": nfd,
            "max_abs_err_vs_fd_rel_to_max_grad": max(errs), "pcg_iterations": patch.iters}


def bench_mf_apply(mf, nel, nb_elem, nrep=20):
    """The matrix-free apply alone at degree 2: DOF-applies/s and HBM roofline (algorithmic bytes:
    the 13 stored geometry values per Gauss point + x read + y written, in the apply's precision)."""
    import torch
    out = {}
    nqp = nb_elem * 27
    stream = torch.cuda.current_stream()
    for name, dtype, sz in (("fp32", torch.float32, 4), ("fp64", torch.float64, 8)):
        x = torch.randn(mf.ndof, dtype=dtype, device="cuda")
        y = torch.empty_like(x)
        mf.apply(x, y)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(nrep):
            mf.apply(x, y)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / nrep
        byts = 13 * sz * nqp + 2 * sz * mf.ndof
        gbs = byts / (ms * 1e-3) / 1e9
        out["apply_" + name] = {"ms": ms, "dof_applies_per_s": mf.ndof / (ms * 1e-3),
                                "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                             "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": byts}}
        log("%d^3 p=2 apply %s: %.3f ms, %.3e DOF-applies/s, %.0f GB/s algorithmic"
            % (nel, name, ms, mf.ndof / (ms * 1e-3), gbs))
        del x, y
    return out


def bench_apply(nel):
    """The matrix-free apply at nel^3, degree 2, and a bounded mixed-precision solve on the same
    operator (apply + fast-diagonalisation preconditioner per iteration)."""
    import torch
    from yeti_amd.geometry import make_box_patch
    from yeti_amd.matfree import MatrixFree
    from yeti_amd.serving import solve_session
    from yeti_amd.solver import free_mask_from_args
    from yeti_amd.assembly import Assembler
    bp = make_box_patch(nel=(nel, nel, nel), p=2)
    asm = Assembler(bp.args)
    F = asm.run_rhs()
    asm.close()
    mask = free_mask_from_args(bp.args)
    mf = MatrixFree(bp.args)
    out = bench_mf_apply(mf, nel, bp.nb_elem)
    for prec in (32, 64):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x, st = solve_session(mf, F, mask, precision=prec, rtol=1e-10, max_iters=200)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        out["solve_%d_200" % prec] = {"seconds": dt, "iterations": st["iters"],
                                      "ms_per_iteration": dt * 1e3 / max(1, st["iters"])}
        log("%d^3 p=2 %s solve, %d iterations: %.2f ms per iteration"
            % (nel, "mixed" if prec == 32 else "fp64", st["iters"], dt * 1e3 / max(1, st["iters"])))
    mf.close()
    return out


def bench_mixed_precision(nel, jacobi_cap):
    """The precision workload (128^3 p=2): the FP64 oracle is the assembled matrix (device COO ->
    CSR of K + K^T) solved in FP64 with the fast-diagonalisation preconditioner to its FP64 floor;
    the mixed solve (FP32 matrix-free apply, FP32 preconditioner and vectors, FP64 residual
    replacement) and the FP64 matrix-free solve are compared with it; the stated FP32 3x3
    block-Jacobi variant runs to a bounded number of iterations."""
    import numpy as np
    import torch
    from yeti_amd.assembly import Assembler
    from yeti_amd.geometry import make_box_patch
    from yeti_amd.matfree import MatrixFree
    from yeti_amd.serving import solve_session
    from yeti_amd.solver import StiffnessCSR, free_mask_from_args
    t0 = time.perf_counter()
    bp = make_box_patch(nel=(nel, nel, nel), p=2)
    log("%d^3 p=2: model generated in %.1f s (%d elements, %d dofs)"
        % (nel, time.perf_counter() - t0, bp.nb_elem, bp.nb_dof_tot))
    out = {}
    asm = Assembler(bp.args)
    F = asm.run_rhs()
    K = torch.empty(asm.nb_data, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    asm.run_values(K)
    csr = StiffnessCSR(asm)
    vals = csr.values(K)
    torch.cuda.synchronize()
    t_csr = time.perf_counter() - t0
    del K
    torch.cuda.empty_cache()
    log("%d^3 p=2: assembled COO values + CSR of K + K^T in %.2f s (nnz %d)" % (nel, t_csr, csr.nnz))
    mask = free_mask_from_args(bp.args)
    maskd = mask.to(torch.float64)
    bnorm = float((F * maskd).norm())
    mf = MatrixFree(bp.args)
    sols = {}

    def run(name, **kw):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x, st = solve_session(mf, F, mask, step_log=lambda s: log("  %s: %d iterations, relres %.2e, %.1f s"
                                                                  % (name, s["iters"], s["relres"],
                                                                     time.perf_counter() - t1)),
                              step_every=20, **kw)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        rel_csr = float(((F - csr.spmv(vals, x)) * maskd).norm()) / bnorm
        log("%d^3 p=2 %s: %.2f s, %d iterations, relres %.2e (assembled matrix %.2e)%s"
            % (nel, name, dt, st["iters"], st["relres"], rel_csr, ", at the FP64 floor" if st["floor"] else ""))
        sols[name] = x
        return {"seconds": dt, "iterations": st["iters"], "replacements": st["updates"], "relres": st["relres"],
                "relres_assembled_csr": rel_csr, "failed": bool(st["failed"]), "stopped_at_fp64_floor": bool(st["floor"])}

    out["fp64_assembled_oracle"] = run("fp64 assembled", precision=64, rtol=1e-10, csr=csr, csr_vals=vals,
                                       max_iters=20000)
    u0 = sols["fp64 assembled"]
    floor = float(np.finfo(np.float64).eps) * float((csr.spmv_abs(vals, u0) * maskd).norm()) / bnorm
    out["fp64_assembled_oracle"]["fp64_floor_eps_absA_absx_over_b"] = floor
    out["mixed"] = run("mixed", precision=32, rtol=1e-10, max_iters=20000)
    out["fp64"] = run("fp64 matrix-free", precision=64, rtol=1e-10, max_iters=20000)
    J0 = float(torch.dot(F, u0))
    for name, key in (("mixed", "mixed"), ("fp64 matrix-free", "fp64")):
        u = sols[name]
        out[key]["displacement_rel_err_vs_assembled_fp64"] = float((u - u0).norm() / u0.norm())
        out[key]["compliance_rel_err_vs_assembled_fp64"] = abs(float(torch.dot(F, u)) - J0) / abs(J0)
    log("%d^3 p=2 errors against the FP64 assembled-matrix solve: mixed |du|/|u| %.2e, |dJ|/J %.2e; "
        "FP64 matrix-free |du|/|u| %.2e; FP64 floor eps||A||u|| / ||b|| %.2e"
        % (nel, out["mixed"]["displacement_rel_err_vs_assembled_fp64"],
           out["mixed"]["compliance_rel_err_vs_assembled_fp64"], out["fp64"]["displacement_rel_err_vs_assembled_fp64"],
           floor))
    out["block_jacobi_fp32"] = run("mixed, 3x3 block-Jacobi", precision=32, precond="jacobi", rtol=1e-10,
                                   max_iters=jacobi_cap)
    out["block_jacobi_fp32"]["max_iterations"] = jacobi_cap
    sols.pop("mixed, 3x3 block-Jacobi")
    xs = {"mixed": sols["mixed"].cpu().numpy(), "fp64": sols["fp64 matrix-free"].cpu().numpy()}
    del sols, u0, vals
    csr.close()
    del csr
    torch.cuda.empty_cache()
    out.update(bench_mf_apply(mf, nel, bp.nb_elem))
    mf.close()
    del mf
    torch.cuda.empty_cache()
    stream = torch.cuda.current_stream()
    # displacement reconstruction (reconstruction.f:22) of the mixed solution, on the device
    import ctypes as C
    from yeti_amd import _lib
    a = bp.args
    nfree, ncp = int(a["nb_dof_free"]), int(a["nb_cp"])
    idf = np.asarray(a["ind_dof_free"])[:nfree].astype(np.int32)
    u_inv = torch.as_tensor(xs["mixed"][idf - 1], device="cuda")
    ind = torch.as_tensor(idf, device="cuda")
    tgt = torch.as_tensor(np.asarray(a["bc_target"]).reshape(-1).astype(np.int32), device="cuda")
    bn = np.ascontiguousarray(np.asarray(a["bc_target_nbelem"]).reshape(-1)[: int(a["nb_bc"])], dtype=np.int32)
    bv = np.ascontiguousarray(np.asarray(a["bc_values"], np.float64).reshape(2, -1, order="F")[:, : bn.size].T)
    sol = torch.empty(3 * ncp, dtype=torch.float64, device="cuda")
    U = torch.empty_like(sol)

    def recon():
        _lib.check(_lib.lib().yeti_reconstruction(
            C.c_void_p(u_inv.data_ptr()), C.c_void_p(ind.data_ptr()), nfree, C.c_void_p(tgt.data_ptr()),
            bn.ctypes.data_as(C.c_void_p), bv.ctypes.data_as(C.c_void_p), int(bn.size), 3, ncp,
            C.c_void_p(sol.data_ptr()), C.c_void_p(U.data_ptr()), C.c_void_p(stream.cuda_stream)))
    recon()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(20):
        recon()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 20
    byts = 12 * nfree + 16 * 3 * ncp     # u_inv + index read, U and sol written
    out["reconstruction"] = {"ms": ms, "dofs": 3 * ncp, "achieved_gbs": byts / (ms * 1e-3) / 1e9,
                             "algorithmic_bytes": byts,
                             "max_abs_diff_vs_solution": float(np.abs(U.cpu().numpy() - xs["mixed"]).max())}
    log("%d^3 p=2 reconstruction: %.3f ms (%.0f GB/s algorithmic)" % (nel, ms, byts / (ms * 1e-3) / 1e9))
    # displacement at 2^20 random parametric points (evaluate_disp.f), host arrays in and out
    from yeti_amd.postproc import evaldisp_points
    pts = np.random.default_rng(1).uniform(0.0, 1.0, (1 << 20, 3))
    solT = U.cpu().numpy().reshape(-1, 3).T
    kw = dict(IEN=a["IEN"], elementsByPatch=a["nb_elem_patch"], Nkv=a["Nkv"], Ukv=a["Ukv"], Nijk=a["Nijk"],
              weight=a["weight"], Jpqr=a["Jpqr"], NNODE=a["NNODE"])
    evaldisp_points(1, pts[:1024], solT, **kw)
    t0 = time.perf_counter()
    evaldisp_points(1, pts, solT, **kw)
    t_ev = time.perf_counter() - t0
    out["evaldisp"] = {"points": pts.shape[0], "seconds_incl_transfers": t_ev,
                       "points_per_s": pts.shape[0] / t_ev}
    log("%d^3 p=2 evaldisp: %d points in %.3f s (host arrays in and out)" % (nel, pts.shape[0], t_ev))
    # VTU output (postproc_vtu.f): vertex fields of every element on the GPU, then the native
    # writer on a 16384-element slice
    from yeti_amd.postproc import svars_solid
    vflags, vref = (True, True, True), (1, 1, 1)
    vargs = (solT, a["COORDS3D"], a["IEN"], a["nb_elem_patch"], a["Nkv"], a["Ukv"], a["Nijk"], a["weight"], a["Jpqr"],
             a["NNODE"], a["MATERIAL_PROPERTIES"])
    t0 = time.perf_counter()
    sv = svars_solid(1, vref, vflags, *vargs)
    t_sv = time.perf_counter() - t0
    import tempfile
    sl = np.ascontiguousarray(sv[:16384])
    with tempfile.TemporaryDirectory() as td:
        vpath = os.path.join(td, "slice.vtu")
        t0 = time.perf_counter()
        _lib.check(_lib.lib().yeti_vtu_write_piece(
            vpath.encode(), sl.ctypes.data_as(C.c_void_p), sl.shape[0], np.array(vref, np.int32).ctypes.data_as(
                C.c_void_p), np.array(vflags, np.int32).ctypes.data_as(C.c_void_p)))
        t_wr = time.perf_counter() - t0
        vbytes = os.path.getsize(vpath)
    out["vtu"] = {"elements": sv.shape[0], "vertices": sv.shape[0] * sv.shape[1],
                  "svars_seconds_incl_transfers": t_sv, "writer_elements": sl.shape[0], "writer_seconds": t_wr,
                  "writer_mb_per_s": vbytes / t_wr / 1e6}
    log("%d^3 p=2 VTU: vertex fields of %d elements in %.3f s (host arrays in and out); writer %.1f MB/s"
        % (nel, sv.shape[0], t_sv, vbytes / t_wr / 1e6))
    del sv, sl
    out["speedup_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["workload"] = ("%d^3 p=2 cantilever (%d dofs). Oracle: the assembled K + K^T (device COO -> CSR) solved in "
                       "FP64, fast-diagonalisation PCG, to its FP64 floor. Mixed: FP32 matrix-free apply, FP32 "
                       "fast-diagonalisation preconditioner and vectors, FP64 residual replacement, to 1e-10; "
                       "errors against the oracle. The stated FP32 3x3 block-Jacobi variant: bounded run."
                       % (nel, bp.nb_dof_tot))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nel", type=int, default=64, help="elements per direction")
    ap.add_argument("--p", type=int, default=3, help="degree")
    ap.add_argument("--cpu-slab", type=int, default=16,
                    help="elements in z of the nel x nel x slab patch timed by the CPU baseline")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU calls (median) after one warm-up")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = every core of this host")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=60.0, help="offered requests per second")
    ap.add_argument("--serve-plain", type=int, default=100,
                    help="requests solved one at a time (plain FP64), from the start of the stream")
    ap.add_argument("--grad-fd", type=int, default=50, help="coordinates checked by central differences")
    ap.add_argument("--jacobi-iters", type=int, default=60000, help="bound on the Jacobi PCG iterations of the assembled-patch solve")
    ap.add_argument("--mp-nel", type=int, default=128, help="elements per direction of the mixed-precision solve")
    ap.add_argument("--mp-jacobi-iters", type=int, default=6000,
                    help="bound on the FP32 block-Jacobi iterations of the mixed-precision section")
    ap.add_argument("--only-assembly", action="store_true")
    ap.add_argument("--sections", default="solve,serving,gradient,mixed",
                    help="comma list of the parts run after the assembly: solve, serving, gradient, mixed, apply "
                         "(the matrix-free apply and bounded solves alone, at --mp-nel)")
    ap.add_argument("--pmc-probe", nargs="?", const="elem", choices=("elem", "mf"),
                    help="run only a 2 GiB device copy, then one element-kernel run (elem) or one FP32 and one "
                         "FP64 matrix-free apply at --mp-nel, degree 2 (mf), for rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: one GPU only (the assembly would shard by contiguous element ranges, "
              "each GPU writing its own slice of the COO with no communication). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from yeti_amd.assembly import Assembler
    from yeti_amd.geometry import make_box_patch

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    p = args.p
    nel = (args.nel, args.nel, args.nel)
    if args.pmc_probe:
        # calibration copy of known size (2 GiB read + 2 GiB written), then one run of the
        # measured kernels; counters per dispatch come from rocprofv3
        a = torch.ones(2 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        if args.pmc_probe == "mf":
            from yeti_amd.matfree import MatrixFree
            bp = make_box_patch(nel=(args.mp_nel,) * 3, p=2)
            mf = MatrixFree(bp.args)
            for dt in (torch.float32, torch.float64):
                x = torch.randn(mf.ndof, dtype=dt, device="cuda")
                mf.apply(x, torch.empty_like(x))
            torch.cuda.synchronize()
            log("pmc probe done: copy of %d bytes, matrix-free apply of %d elements" % (a.numel(), bp.nb_elem))
            mf.close()
            return
        bp = make_box_patch(nel=nel, p=p)
        asm = Assembler(bp.f2py_args())
        K, r, c = asm.alloc_coo(True)
        asm.run_matrix(K, r, c)
        torch.cuda.synchronize()
        log("pmc probe done: copy of %d bytes, %d elements" % (a.numel(), bp.nb_elem))
        return
    t0 = time.perf_counter()
    bp = make_box_patch(nel=nel, p=p)
    log("generated %s p=%d in %.1f s: %d elements, %d dofs, nb_data %d"
        % (nel, p, time.perf_counter() - t0, bp.nb_elem, bp.nb_dof_tot, bp.nb_data))
    t0 = time.perf_counter()
    asm = Assembler(bp.f2py_args())
    log("resident set-up %.1f s" % (time.perf_counter() - t0))
    K, r, c = asm.alloc_coo(True)
    F = torch.empty(asm.nb_dof_tot, dtype=torch.float64, device="cuda")
    log("COO buffers: %.1f GB" % (asm.nb_data * 16 / 1e9))

    for _ in range(args.warmup):
        asm.run_coo(K, r, c, F)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        asm.run_coo(K, r, c, F)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    value = bp.nb_elem / (dt / args.steps)
    log("timed %d steps: %.3f ms/step, %.3e elem/s" % (args.steps, ms, value))

    # ---- roofline of the dominant kernel (element kernel with indices) -------------------
    n = (p + 1) ** 3
    nq = (p + 1) ** 3
    e_ent = 9 * n * (n - 1) // 2 + 6 * n
    bytes_elem = 16 * e_ent + 12 * n + 12            # values + row + col out; IEN + weights + spans in
    flops_elem = 9 * n * (n + 1) * nq
    stream = torch.cuda.current_stream()
    nrep = max(3, args.steps)
    asm.set_timing(True)
    launch_ms = []
    for _ in range(nrep):
        asm.run_matrix(K, r, c)
        launch_ms += asm.kernel_times()
    nlaunch = len(asm.kernel_times())
    kern_ms = sum(launch_ms) / len(launch_ms)                    # average launch duration
    elems_per_launch = bp.nb_elem / nlaunch
    bytes_launch = elems_per_launch * bytes_elem + 24 * bp.nb_cp / nlaunch
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9
    launch_ms = []
    for _ in range(nrep):
        asm.run_values(K)
        launch_ms += asm.kernel_times()
    val_ms = sum(launch_ms) / len(launch_ms)
    asm.set_timing(False)
    tflops = elems_per_launch * flops_elem / (val_ms * 1e-3) / 1e12
    log("element kernel: %d launches of %d elements, %.3f ms each (%.0f GB/s); values-only %.3f ms (%.1f TFLOP/s FP64)"
        % (nlaunch, elems_per_launch, kern_ms, achieved, val_ms, tflops))
    traffic = read_traffic(os.path.join(ROOT, "profiles", "elem_stiffness_64cube_p%d_traffic.md" % p))
    # consistent mass matrix of the same model in the same COO layout (compute_CMASSMatrix.f),
    # written into the same buffers: HBM-bound (values + indices written)
    asm.run_mass(K, r, c)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        asm.run_mass(K, r, c)
    torch.cuda.synchronize()
    mass_ms = (time.perf_counter() - t0) / 3 * 1e3
    mass_bytes = bp.nb_elem * bytes_elem + 24 * bp.nb_cp
    log("mass matrix (COO values + indices): %.2f ms, %.0f GB/s" % (mass_ms, mass_bytes / (mass_ms * 1e-3) / 1e9))
    del K, r, c, F
    torch.cuda.empty_cache()
    # natural-frequency shape gradient (gradvibration_fullAN.f) of 4 synthetic modes on the same
    # model (the first call also builds the CP -> element-node map)
    nmode = 4
    vect = torch.randn((nmode, 3, bp.nb_cp), dtype=torch.float64, device="cuda")
    lam = np.linspace(1e6, 4e7, nmode)
    asm.run_vib_grad(vect, lam)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        gv = asm.run_vib_grad(vect, lam)
    torch.cuda.synchronize()
    vib_ms = (time.perf_counter() - t0) / 3 * 1e3
    log("vibration gradient (%d modes): %.2f ms, %.3e elem/s" % (nmode, vib_ms, bp.nb_elem / (vib_ms * 1e-3)))
    del vect, gv
    # element-wise strain-energy gradient -d(u^T K u)/dX (the compliance gradient's stiffness term
    # for models the matrix-free operator does not cover) and the volume gradient, same model
    u = torch.randn(3 * bp.nb_cp, dtype=torch.float64, device="cuda")
    asm.strain_energy_grad(u)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        gs = asm.strain_energy_grad(u)
    torch.cuda.synchronize()
    seg_ms = (time.perf_counter() - t0) / 3 * 1e3
    t0 = time.perf_counter()
    for _ in range(3):
        vol, gvol = asm.volume_grad()
    torch.cuda.synchronize()
    vol_ms = (time.perf_counter() - t0) / 3 * 1e3
    log("strain-energy gradient: %.2f ms; volume + gradient: %.2f ms" % (seg_ms, vol_ms))
    del u, gs, vol, gvol
    torch.cuda.empty_cache()

    def grad_work(ms, nf):
        """Algorithmic work of one gradient call with nf fields (modes) against the HBM and FP64
        peaks. Per element: IEN, weights, coordinates, spans and the fields' nodal values read,
        per-node partials and the mass written, then gathered per CP; flops: R and dR/dx per
        (point, node) (~26), u and du/dx per (point, field) (24 per node), the tensor contraction
        per (point, node, field) (18)."""
        n = (p + 1) ** 3
        nq = (p + 1) ** 3
        byts = bp.nb_elem * (4 * n + 8 * n + 24 * n + 12 + nf * (24 * n + 24 * n + 8)) + nf * 24 * bp.nb_cp
        flops = bp.nb_elem * nq * n * (26 + nf * (24 + 18))
        s = ms * 1e-3
        return {"algorithmic_bytes": byts, "achieved_gbs": byts / s / 1e9, "hbm_frac": byts / s / 1e9 / HBM_PEAK_GBS,
                "flops": flops, "achieved_tflops": flops / s / 1e12, "fp64_frac": flops / s / 1e12 / 78.6,
                "bound": "latency (one workgroup per element, 5 barriers per Gauss-point chunk)"}

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        log("cpu baseline: %dx%dx%d elements with %d threads ..." % (args.nel, args.nel, args.cpu_slab, threads))
        cpu = cpu_baseline(args.nel, p, args.cpu_slab, args.cpu_runs, threads)
        log("cpu baseline: %.3e elem/s (%s); 16x4x4 p=2: %.3e elem/s"
            % (cpu["value"], cpu["kind"], cpu["cantilever_16x4x4_p2"]["value"]))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "FP64 COO stiffness (values + int32 row/col) + traction RHS of a 10x1x1 "
                        "cantilever, %dx%dx%d NURBS elements, degree %d, %d^3 Gauss points, "
                        "jittered control points and weights" % (nel + (p, p + 1)),
            "elements": bp.nb_elem, "dofs": bp.nb_dof_tot, "coo_entries": bp.nb_data,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "elem_stiffness_kernel<%d,%d,true>" % (p, p + 1),
            "kernel_ms": kern_ms, "launches_per_step": nlaunch, "elements_per_launch": elems_per_launch,
            "algorithmic_bytes_per_launch": bytes_launch,
        },
        "mfma_values_only": {
            "kernel": "elem_stiffness_kernel<%d,%d,false>" % (p, p + 1), "kernel_ms": val_ms,
            "achieved": tflops, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
            "frac": tflops / FP64_PEAK_TFS, "flops_per_launch": elems_per_launch * flops_elem,
        },
        "cpu_baseline": cpu,
        "mass_matrix": {"ms": mass_ms, "elements_per_s": bp.nb_elem / (mass_ms * 1e-3),
                        "roofline": {"bound": "hbm", "achieved": mass_bytes / (mass_ms * 1e-3) / 1e9,
                                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": mass_bytes / (mass_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}},
        "vibration_gradient": {"ms": vib_ms, "modes": nmode, "elements_per_s": bp.nb_elem / (vib_ms * 1e-3),
                               "work": grad_work(vib_ms, nmode)},
        "strain_energy_gradient": {"ms": seg_ms, "elements_per_s": bp.nb_elem / (seg_ms * 1e-3),
                                   "work": grad_work(seg_ms, 1)},
        "volume_gradient": {"ms": vol_ms, "elements_per_s": bp.nb_elem / (vol_ms * 1e-3), "work": grad_work(vol_ms, 1)},
    }
    sections = set() if args.only_assembly else set(args.sections.split(","))
    if "solve" in sections:
        out["solve_64cube_p3"] = bench_solve_64cube(args.nel, p, args.jacobi_iters)
    if "serving" in sections:
        out["serving"] = bench_serving(args.serve_requests, args.serve_rate, 20261202, 200 << 30, args.serve_plain)
    if "gradient" in sections:
        out["gradient"] = bench_gradient(args.grad_fd)
    if "apply" in sections:
        out["matfree_apply"] = bench_apply(args.mp_nel)
    if "mixed" in sections:
        out["mixed_precision"] = bench_mixed_precision(args.mp_nel, args.mp_jacobi_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
