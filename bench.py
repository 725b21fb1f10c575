#!/usr/bin/env python3
"""Throughput of YETI's 3D NURBS solid (U1) FP64 assembly on one MI355X.

Workload (one step): the full drop-in output of sys_linmat_lindef_static
(build_system_elemStorage.f90:24-315) for a 10x1x1 cantilever meshed with 64x64x64 degree-3
NURBS elements (262,144 elements, 902,289 dofs, 4.86e9 COO entries): every element's
upper-triangle stiffness blocks (FP64 values + int32 row/col ids, 77.7 GB) and the traction
right-hand side F, written to HBM. Inputs (control points, connectivity, knots, weights) are
resident in HBM before the timed region starts.

Prints one JSON line on stdout; progress goes to stderr. ``roofline`` describes the dominant
kernel (the element kernel with index emission), timed with HIP events on the stream it runs
on; ``cpu_baseline`` is YETI's own OpenMP Fortran routine (oracle/_ref) on a bounded sample of
the same workload, on this host's cores.
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

METRIC = "assembled elem/s + served solves/s, p99 latency vs plain FP64 at rel err ≤1e-8"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(p, sample_nel, threads):
    """YETI's sys_linmat_lindef_static_omp (compiled reference) on a sample patch; falls back to
    the C restatement (oracle/liboracle.so) when the reference library is absent."""
    from oracle import oracle as O
    from yeti_amd.geometry import make_box_patch
    bp = make_box_patch(nel=sample_nel, p=p, seed=20261202)
    args = bp.f2py_args()
    kind = "reference" if O.ref_available() else "port"
    run = (lambda: O.ref_sys_linmat_omp(args, num_threads=threads)) if kind == "reference" else \
        (lambda: O.oracle_sys_linmat(args, nthreads=threads))
    t0 = time.perf_counter()
    run()
    dt = time.perf_counter() - t0
    ne = bp.nb_elem
    return {"value": ne / dt, "unit": "elem/s", "cores": threads, "kind": kind,
            "sample": "%dx%dx%d elements p=%d (%d elements, same generator), one call of "
                      "sys_linmat_lindef_static_omp incl. output allocation, %.1f s"
                      % (sample_nel[0], sample_nel[1], sample_nel[2], p, ne, dt)}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nel", type=int, default=64, help="elements per direction")
    ap.add_argument("--p", type=int, default=3, help="degree")
    ap.add_argument("--cpu-sample", default="32,16,16", help="elements per direction of the CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the assembly shards by element "
              "ranges with no communication; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
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
    bytes_launch = bp.nb_elem * bytes_elem + 24 * bp.nb_cp
    flops_elem = 9 * n * (n + 1) * nq
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nrep = max(3, args.steps)
    ev0.record(stream)
    for _ in range(nrep):
        asm.run_matrix(K, r, c)
    ev1.record(stream)
    torch.cuda.synchronize()
    kern_ms = ev0.elapsed_time(ev1) / nrep
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9
    # values-only launch: the MFMA-bound variant (indices are static and emitted once)
    ev0.record(stream)
    for _ in range(nrep):
        asm.run_values(K)
    ev1.record(stream)
    torch.cuda.synchronize()
    val_ms = ev0.elapsed_time(ev1) / nrep
    tflops = bp.nb_elem * flops_elem / (val_ms * 1e-3) / 1e12
    log("element kernel: %.3f ms (%.0f GB/s); values-only %.3f ms (%.1f TFLOP/s FP64)"
        % (kern_ms, achieved, val_ms, tflops))

    cpu = None
    if not args.no_cpu_baseline:
        sample = tuple(int(x) for x in args.cpu_sample.split(","))
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log("cpu baseline on %s with %d threads ..." % (sample, threads))
        cpu = cpu_baseline(p, sample, threads)
        log("cpu baseline: %.3e elem/s (%s)" % (cpu["value"], cpu["kind"]))

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
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "elem_stiffness_kernel<%d,%d,true>" % (p, p + 1),
            "kernel_ms": kern_ms, "algorithmic_bytes_per_launch": bytes_launch,
        },
        "mfma_values_only": {
            "kernel": "elem_stiffness_kernel<%d,%d,false>" % (p, p + 1), "kernel_ms": val_ms,
            "achieved": tflops, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
            "frac": tflops / FP64_PEAK_TFS, "flops_per_launch": bp.nb_elem * flops_elem,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
