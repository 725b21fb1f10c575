"""Device-resident request loop: assembly + solve requests for NURBS solid patches of different
sizes and degrees, admitted against an HBM budget, solved with graph-captured, device-controlled
solve steps, reporting per-request latency next to throughput.

No reference counterpart exists: YETI builds and solves one model per call
(build_system_elemStorage.f90:24-315, then scipy or iga_mf_cg_3d in its callers). What a request
computes is the reference's pipeline: F from the load description
(build_system_elemStorage.f90:143-152,295-303) and u from K[free][:, free] u = F[free]
(benchs/volume_force/test_volume_force_3d.py:67-79), with K applied matrix-free.

Memory: one device arena of `budget_bytes` is allocated when the server starts. A request is
admitted (first come, first served) when its footprint -- measured exactly by the set-up
routines' counting mode, plus the solve session and the request's vectors -- fits in a free
range of the arena; its handles are created inside that range (no hipMalloc/hipFree while
serving) and the range is returned when the request completes.

Solves: ``precision="mixed"`` is PCG (fast-diagonalisation or block-Jacobi preconditioner)
with FP32 vectors, operator and preconditioner (FP64 dots) whose FP64 solution and true residual are refreshed by residual
replacement with the FP64 apply each time the FP32 residual has fallen by ``delta``
(refinement that keeps the Krylov space); ``precision="fp64"`` runs the same PCG in FP64.
Both stop on the FP64 residual. Each step (``chunk`` iterations) is one hipGraph launch per
request; steps of different requests run on a pool of streams and never change each other's
iterates, so a request's solution is bit-identical alone or in a mix.
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._abi import YETI_LAYOUT_PACKED, pack_system_inputs
from .geometry import make_box_patch

ALIGN = 256
NRED = 1024
BATCH_MAX_MEMBERS = 256   # members of one batch (the library's table size)
STATUS_BYTES = 8 * 4 + 16 * 8     # yeti_ir_status


def _align(n):
    return (int(n) + ALIGN - 1) // ALIGN * ALIGN


# ------------------------------------------------------------------------------------------
class ArenaAllocator:
    """First-fit allocator of byte ranges of [0, size), ALIGN-aligned, with coalescing frees.
    Host bookkeeping only; the device memory is one buffer allocated up front."""

    def __init__(self, size):
        self.size = int(size)
        self.free = [(0, self.size)]      # sorted, disjoint, non-adjacent (offset, length)
        self.used = 0
        self.peak = 0

    def alloc(self, nbytes):
        n = _align(max(int(nbytes), 1))
        for i, (off, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + n, ln - n)
                self.used += n
                self.peak = max(self.peak, self.used)
                return off
        return None

    def release(self, off, nbytes):
        n = _align(max(int(nbytes), 1))
        if off < 0 or off + n > self.size:
            raise ValueError("range outside the arena")
        i = 0
        while i < len(self.free) and self.free[i][0] < off:
            i += 1
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] > off:
            raise ValueError("double free")
        if i < len(self.free) and off + n > self.free[i][0]:
            raise ValueError("double free")
        self.free.insert(i, (off, n))
        # coalesce with the right, then the left neighbour
        if i + 1 < len(self.free) and off + n == self.free[i + 1][0]:
            self.free[i] = (off, n + self.free[i + 1][1])
            self.free.pop(i + 1)
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == off:
            self.free[i - 1] = (self.free[i - 1][0], self.free[i - 1][1] + self.free[i][1])
            self.free.pop(i)
        self.used -= n


def ir_session_bytes(ncp_dims, precision, precond="fdm"):
    """Device bytes of yeti_ir_create (mirrors its allocations, each ALIGN-aligned; the fast
    diagonalisation part is the bound of fdm_bytes_bound, no face removed)."""
    nb_cp = int(np.prod(ncp_dims))
    n = 3 * nb_cp
    w = 8 if precision == 64 else 4
    sizes = [9 * nb_cp * 8, n * 8, n * 8, NRED * 8, 16 * 8, 8 * 4, 9 * nb_cp * w] + [n * w] * 5
    total = sum(_align(s) for s in sizes)
    if precond == "fdm":
        al = lambda b: _align(b) + ALIGN   # noqa: E731
        mats = 3 * sum(int(d) * int(d) for d in ncp_dims) * w
        total += 2 * al(mats) + al(3 * nb_cp * w) + 2 * al(3 * nb_cp * w)
    return total


# ------------------------------------------------------------------------------------------
@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    nel: tuple
    p: int
    E: float
    nu: float
    traction: float
    seed: int
    args: dict = field(default=None, repr=False)

    def model(self):
        if self.args is None:
            self.args = make_box_patch(nel=self.nel, p=self.p, E=self.E, nu=self.nu, traction=self.traction,
                                       seed=self.seed).args
        return self.args

    @property
    def ncp_dims(self):
        return tuple(self.nel[d] + self.p for d in range(3))

    @property
    def nb_cp(self):
        return (self.nel[0] + self.p) * (self.nel[1] + self.p) * (self.nel[2] + self.p)


def request_stream(n=2000, rate=200.0, seed=20261202, nel_range=(4, 48), degrees=(2, 3, 4)):
    """Poisson arrivals at `rate` per second; elements per direction round(exp(U[ln a, ln b]));
    degree uniform over `degrees`; E ~ U[1e3, 1e5], nu ~ U[0.2, 0.4], traction ~ U[0.1, 10]
    (downward, on the face x = 10 of a 10 x 1 x 1 box clamped at x = 0; interior control
    points jittered by 5 % of an element)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    lo, hi = math.log(nel_range[0]), math.log(nel_range[1])
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        nel = tuple(int(round(math.exp(rng.uniform(lo, hi)))) for _ in range(3))
        p = int(rng.choice(degrees))
        E = float(rng.uniform(1e3, 1e5))
        nu = float(rng.uniform(0.2, 0.4))
        tr = -float(rng.uniform(0.1, 10.0))
        out.append(Request(i, t, nel, p, E, nu, tr, int(rng.integers(1 << 31))))
    return out


def latency_summary(lat):
    a = np.sort(np.asarray(lat, dtype=np.float64))
    if a.size == 0:
        return {"p50": None, "p99": None, "mean": None, "max": None}
    def pct(q):
        # nearest-rank percentile
        k = max(0, min(a.size - 1, int(math.ceil(q / 100.0 * a.size)) - 1))
        return float(a[k])
    return {"p50": pct(50), "p99": pct(99), "mean": float(a.mean()), "max": float(a[-1])}


def solve_session(mf, F, mask, precision=64, precond="fdm", rtol=1e-10, delta=0.1, chunk=32, max_iters=100000,
                  max_updates=2000, csr=None, csr_vals=None, step_log=None, step_every=0):
    """One graph-stepped solve session (the serving loop's solver) on a resident MatrixFree
    operator ``mf``: K[free][:, free] x = F[free] to rtol on the FP64 residual. precision 64 or 32
    (FP32 vectors, operator and preconditioner with FP64 residual replacement). With ``csr`` (a
    StiffnessCSR) and its values, the FP64 products use the assembled matrix: the FP64 solve of
    the assembled system. Returns (x, status dict)."""
    import torch
    L = _lib.lib()
    n = int(F.numel())
    ncp = n // 3
    used = C.c_size_t(0)
    h = C.c_void_p()
    mask_p = C.c_void_p(mask.data_ptr())
    _lib.check(L.yeti_ir_create(mf._h, mask_p, int(precision), 1 if precond == "fdm" else 0, float(rtol), float(delta),
                                int(chunk), int(max_iters), int(max_updates), None, 0, None, C.byref(h),
                                C.byref(used)))
    arena = torch.empty(int(used.value) + (1 << 20), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    torch.cuda.synchronize()
    _lib.check(L.yeti_ir_create(mf._h, mask_p, int(precision), 1 if precond == "fdm" else 0, float(rtol), float(delta),
                                int(chunk), int(max_iters), int(max_updates), C.c_void_p(arena.data_ptr()),
                                int(arena.numel()), sh, C.byref(h), C.byref(used)))
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    status = torch.empty(STATUS_BYTES, dtype=torch.uint8, pin_memory=True)
    try:
        if csr is not None:
            _lib.check(L.yeti_ir_set_csr(h, csr._h, C.c_void_p(csr_vals.data_ptr())))
        _lib.check(L.yeti_ir_begin(h, C.c_void_p(F.data_ptr()), C.c_void_p(x.data_ptr()), sh))
        steps = 0
        while True:
            _lib.check(L.yeti_ir_step(h, 1, sh))
            _lib.check(L.yeti_ir_status_async(h, C.c_void_p(status.data_ptr()), sh))
            stream.synchronize()
            steps += 1
            buf = status.numpy()
            fl = np.frombuffer(buf[:32].tobytes(), dtype=np.int32)
            sc = np.frombuffer(buf[32:].tobytes(), dtype=np.float64)
            st = {"done": int(fl[3]), "failed": int(fl[4]), "floor": int(fl[6]), "iters": int(fl[5]),
                  "updates": int(fl[7]), "relres": math.sqrt(sc[5] / sc[4]) if sc[4] > 0 else 0.0}
            if step_log and step_every and steps % step_every == 0:
                step_log(st)
            if st["done"]:
                return x, st
    finally:
        stream.synchronize()
        L.yeti_ir_destroy(h)
        del arena


# ------------------------------------------------------------------------------------------
class _Active:
    __slots__ = ("req", "off", "nbytes", "asm", "mf", "ir", "F", "x", "mask", "stream", "event", "status",
                 "t_admit", "steps", "keep", "in_batch")


class SolveServer:
    """Runs a list of requests against one arena of `budget_bytes`.

    Each active request is a solve session on one of `nstreams` streams; HIP spreads streams over
    GPU_MAX_HW_QUEUES hardware queues (default 4), so a serving process should set that variable
    to about `nstreams` before the HIP runtime starts for the sessions to overlap."""

    def __init__(self, budget_bytes, precision="mixed", nstreams=16, chunk=16, rtol=1e-8, delta=0.1,
                 max_iters=100000, max_updates=400, use_graph=True, prep_threads=2, precond="fdm", batched=False,
                 task_capacity=1 << 20):
        import torch
        _lib.require_device()
        self.batched = bool(batched)
        if self.batched and precond != "fdm":
            raise ValueError("batched serving runs the fast-diagonalisation preconditioner")
        if precision not in ("mixed", "fp64"):
            raise ValueError("precision must be 'mixed' or 'fp64'")
        self.torch = torch
        self.inner = 32 if precision == "mixed" else 64
        self.precision = precision
        self.chunk, self.rtol, self.delta = int(chunk), float(rtol), float(delta)
        self.max_iters, self.max_updates = int(max_iters), int(max_updates)
        self.use_graph = bool(use_graph)
        if precond not in ("fdm", "jacobi"):
            raise ValueError("precond must be 'fdm' or 'jacobi'")
        self.precond = precond
        self.budget = int(budget_bytes)
        self.arena_buf = torch.empty(self.budget, dtype=torch.uint8, device="cuda")
        self.base = self.arena_buf.data_ptr()
        self.alloc = ArenaAllocator(self.budget)
        self.streams = [torch.cuda.Stream() for _ in range(int(nstreams))]
        self.prep_threads = int(prep_threads)
        self.batch = None
        if self.batched:
            h = C.c_void_p()
            _lib.check(_lib.lib().yeti_batch_create(self.inner, self.chunk, int(task_capacity), C.byref(h)))
            self.batch = h
            self.batch_status = torch.empty((BATCH_MAX_MEMBERS, STATUS_BYTES), dtype=torch.uint8, pin_memory=True)
            self.batch_event = torch.cuda.Event()

    # -- footprint ---------------------------------------------------------------------------
    def footprint(self, req, inputs):
        L = _lib.lib()
        used_a, used_m = C.c_size_t(0), C.c_size_t(0)
        h = C.c_void_p()
        _lib.check(L.yeti_asm_create_in(C.byref(inputs), YETI_LAYOUT_PACKED, None, 0, None, C.byref(h),
                                        C.byref(used_a)))
        _lib.check(L.yeti_mf_create_in(C.byref(inputs), 1 | 2 if self.inner == 32 else 1, None, 0, None, C.byref(h),
                                       C.byref(used_m)))
        n = 3 * req.nb_cp
        parts = [_align(used_a.value) + ALIGN, _align(used_m.value) + ALIGN,
                 ir_session_bytes(req.ncp_dims, self.inner, self.precond) + ALIGN,
                 _align(8 * n), _align(8 * n), _align(n)]
        return parts

    # -- admission -----------------------------------------------------------------------
    def _admit(self, req, stream, keep, join=True):
        torch = self.torch
        L = _lib.lib()
        args = req.model()
        inputs, hold = pack_system_inputs(args)
        parts = self.footprint(req, inputs)
        total = sum(parts)
        off = self.alloc.alloc(total)
        if off is None:
            return None
        a = _Active()
        a.req, a.off, a.nbytes, a.stream, a.keep = req, off, total, stream, keep
        a.t_admit = time.perf_counter()
        a.steps = 0
        sh = C.c_void_p(stream.cuda_stream)
        o = off
        ptr = lambda k: C.c_void_p(self.base + k)   # noqa: E731
        used = C.c_size_t(0)
        try:
            h = C.c_void_p()
            _lib.check(L.yeti_asm_create_in(C.byref(inputs), YETI_LAYOUT_PACKED, ptr(o), parts[0], sh, C.byref(h),
                                            C.byref(used)))
            a.asm = h
            o += parts[0]
            h = C.c_void_p()
            _lib.check(L.yeti_mf_create_in(C.byref(inputs), 1 | 2 if self.inner == 32 else 1, ptr(o), parts[1], sh,
                                           C.byref(h), C.byref(used)))
            a.mf = h
            o += parts[1]
            o_ir = o
            o += parts[2]
            n = 3 * req.nb_cp
            a.F = self.arena_buf[o:o + 8 * n].view(torch.float64)
            o += parts[3]
            a.x = self.arena_buf[o:o + 8 * n].view(torch.float64)
            o += parts[4]
            a.mask = self.arena_buf[o:o + n]
            idf = np.asarray(args["ind_dof_free"]).astype(np.int64)
            m = np.zeros(n, np.uint8)
            m[idf[: int(args["nb_dof_free"])] - 1] = 1
            with torch.cuda.stream(stream):
                a.mask.copy_(torch.from_numpy(m), non_blocking=False)
            _lib.check(L.yeti_asm_run_rhs(a.asm, C.c_void_p(a.F.data_ptr()), sh))
            h = C.c_void_p()
            _lib.check(L.yeti_ir_create(a.mf, C.c_void_p(a.mask.data_ptr()), self.inner,
                                        1 if self.precond == "fdm" else 0, self.rtol, self.delta,
                                        self.chunk, self.max_iters, self.max_updates, ptr(o_ir), parts[2], sh,
                                        C.byref(h), C.byref(used)))
            a.ir = h
            _lib.check(L.yeti_ir_begin(a.ir, C.c_void_p(a.F.data_ptr()), C.c_void_p(a.x.data_ptr()), sh))
        except Exception:
            self._release(a)
            raise
        if self.batched:
            if join:
                _lib.check(L.yeti_batch_add(self.batch, a.ir))
                a.in_batch = True
            return a
        a.status = torch.empty(STATUS_BYTES, dtype=torch.uint8, pin_memory=True)
        a.event = torch.cuda.Event()
        self._step(a)
        return a

    def _step(self, a):
        L = _lib.lib()
        sh = C.c_void_p(a.stream.cuda_stream)
        _lib.check(L.yeti_ir_step(a.ir, 1 if self.use_graph else 0, sh))
        _lib.check(L.yeti_ir_status_async(a.ir, C.c_void_p(a.status.data_ptr()), sh))
        a.event.record(a.stream)
        a.steps += 1

    @staticmethod
    def _parse_status(buf):
        fl = np.frombuffer(buf[:32].tobytes(), dtype=np.int32)
        sc = np.frombuffer(buf[32:].tobytes(), dtype=np.float64)
        return {"done": int(fl[3]), "failed": int(fl[4]), "floor": int(fl[6]), "iters": int(fl[5]),
                "updates": int(fl[7]), "relres": math.sqrt(sc[5] / sc[4]) if sc[4] > 0 else 0.0}

    def _read_status(self, a):
        buf = a.status.numpy()
        fl = np.frombuffer(buf[:32].tobytes(), dtype=np.int32)
        sc = np.frombuffer(buf[32:].tobytes(), dtype=np.float64)
        return {"done": int(fl[3]), "failed": int(fl[4]), "floor": int(fl[6]), "iters": int(fl[5]),
                "updates": int(fl[7]), "relres": math.sqrt(sc[5] / sc[4]) if sc[4] > 0 else 0.0}

    def _release(self, a):
        L = _lib.lib()
        if self.batched and getattr(a, "ir", None) is not None and getattr(a, "in_batch", False):
            _lib.check(L.yeti_batch_remove(self.batch, a.ir))
            a.in_batch = False
        # the handles live in the arena: deleting them frees no device memory
        if getattr(a, "ir", None) is not None:
            L.yeti_ir_destroy(a.ir)
            a.ir = None
        if getattr(a, "mf", None) is not None:
            L.yeti_mf_destroy(a.mf)
            a.mf = None
        if getattr(a, "asm", None) is not None:
            L.yeti_asm_destroy(a.asm)
            a.asm = None
        self.alloc.release(a.off, a.nbytes)

    # -- the loop ------------------------------------------------------------------------
    def run(self, requests, keep_solutions=False, realtime=True, progress=None, step_log=None, step_every=0):
        """Serve `requests` (sorted by arrival). realtime=False admits every request at t=0
        (closed-loop throughput). Returns a report dict; with keep_solutions, report["x"][rid]
        holds each solution on the host. progress(done, total) is called as requests complete;
        step_log(rid, status) every `step_every` steps of a request."""
        if self.batched:
            return self._run_batched(requests, keep_solutions, realtime, progress)
        torch = self.torch
        reqs = sorted(requests, key=lambda r: r.arrival)
        ready = {}
        lock = threading.Lock()
        stop = [False]

        def prep(chunk):
            for r in chunk:
                if stop[0]:
                    return
                r.model()
                with lock:
                    ready[r.rid] = True

        # clients prepare their models ahead of time, in arrival order
        threads = []
        for k in range(max(1, self.prep_threads)):
            th = threading.Thread(target=prep, args=(reqs[k::max(1, self.prep_threads)],), daemon=True)
            th.start()
            threads.append(th)
        pending = deque(reqs)
        active = []
        done = {}
        xs = {}
        t0 = time.perf_counter()
        si = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            while pending or active:
                now = time.perf_counter() - t0
                progressed = False
                # admission, first come first served
                while pending and (not realtime or pending[0].arrival <= now):
                    r = pending[0]
                    with lock:
                        ok = ready.get(r.rid, False)
                    if not ok:
                        break
                    a = self._admit(r, self.streams[si % len(self.streams)], keep_solutions)
                    if a is None:
                        if not active:
                            raise RuntimeError("request %d needs more than the whole budget" % r.rid)
                        break
                    si += 1
                    pending.popleft()
                    active.append(a)
                    progressed = True
                # progress
                still = []
                for a in active:
                    if not a.event.query():
                        still.append(a)
                        continue
                    st = self._read_status(a)
                    if step_log and step_every and a.steps % step_every == 0:
                        step_log(a.req.rid, st)
                    if st["done"]:
                        t_end = time.perf_counter() - t0
                        arr = a.req.arrival if realtime else 0.0
                        done[a.req.rid] = dict(latency=t_end - arr, finish=t_end, steps=a.steps, p=a.req.p,
                                               nel=a.req.nel, ndof=3 * a.req.nb_cp, **st)
                        if a.keep:
                            xs[a.req.rid] = a.x.cpu().numpy().copy()
                        self._release(a)
                        if progress:
                            progress(len(done), len(reqs))
                    else:
                        self._step(a)
                        still.append(a)
                    progressed = True
                active = still
                if not progressed:
                    time.sleep(20e-6)
            torch.cuda.synchronize()
            makespan = time.perf_counter() - t0
        finally:
            stop[0] = True
            for a in active:
                try:
                    a.event.synchronize()
                    self._release(a)
                except Exception:
                    pass
        lat = [d["latency"] for d in done.values()]
        rep = {"completed": len(done), "makespan_s": makespan,
               "solves_per_s": len(done) / makespan if makespan > 0 else None,
               "latency_s": latency_summary(lat), "failed": sum(1 for d in done.values() if d["failed"]),
               "at_fp64_floor": sum(1 for d in done.values() if d["floor"]),
               "peak_arena_bytes": self.alloc.peak, "budget_bytes": self.budget, "requests": done}
        if keep_solutions:
            rep["x"] = xs
        return rep

    def _run_batched(self, requests, keep_solutions, realtime, progress):
        """The loop of a batched server: admitted requests join one batch of sessions; each step
        advances every member by `chunk` iterations with shared launches (one graph) and copies
        the whole batch's status once. While a step runs on the GPU the host admits arrived
        requests on a second stream (set-up, right-hand side, first direction); they join the
        batch after the step, when completed members leave it."""
        torch = self.torch
        L = _lib.lib()
        reqs = sorted(requests, key=lambda r: r.arrival)
        ready = {}
        lock = threading.Lock()
        stop = [False]

        def prep(chunk):
            for r in chunk:
                if stop[0]:
                    return
                r.model()
                with lock:
                    ready[r.rid] = True

        threads = []
        for k in range(max(1, self.prep_threads)):
            th = threading.Thread(target=prep, args=(reqs[k::max(1, self.prep_threads)],), daemon=True)
            th.start()
            threads.append(th)
        stream = self.streams[0]
        adm_stream = self.streams[1] if len(self.streams) > 1 else torch.cuda.Stream()
        sh = C.c_void_p(stream.cuda_stream)
        pending = deque(reqs)
        members = []            # in batch order
        joining = []            # admitted during the running step
        done = {}
        xs = {}
        t_admit = t_wait = 0.0
        nsteps = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()

        def admit_arrived(deadline_event=None, limit=BATCH_MAX_MEMBERS):
            # admit while requests have arrived, fit, and (when a step runs) until it ends
            nonlocal t_admit
            n_admitted = 0
            while pending and (not realtime or pending[0].arrival <= time.perf_counter() - t0) and \
                    len(members) + len(joining) < BATCH_MAX_MEMBERS and n_admitted < limit:
                if deadline_event is not None and deadline_event.query() and (members or joining):
                    return
                n_admitted += 1
                r = pending[0]
                with lock:
                    ok = ready.get(r.rid, False)
                if not ok:
                    return
                ta = time.perf_counter()
                a = self._admit(r, adm_stream, keep_solutions, join=False)
                t_admit += time.perf_counter() - ta
                if a is None:
                    if not members and not joining:
                        raise RuntimeError("request %d needs more than the whole budget" % r.rid)
                    return
                a.steps = 0
                pending.popleft()
                joining.append(a)

        try:
            while pending or members or joining:
                if joining:
                    adm_stream.synchronize()
                    for a in joining:
                        _lib.check(L.yeti_batch_add(self.batch, a.ir))
                        a.in_batch = True
                    members.extend(joining)
                    joining = []
                if not members:
                    # a first group, then the rest while the GPU steps
                    admit_arrived(limit=8)
                    if not joining:
                        if pending and realtime:
                            time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
                        else:
                            time.sleep(20e-6)
                    continue
                tw = time.perf_counter()
                _lib.check(L.yeti_batch_step(self.batch, 1 if self.use_graph else 0, sh))
                _lib.check(L.yeti_batch_status_async(self.batch, C.c_void_p(self.batch_status.data_ptr()), sh))
                self.batch_event.record(stream)
                admit_arrived(self.batch_event)
                self.batch_event.synchronize()
                t_wait += time.perf_counter() - tw
                nsteps += 1
                buf = self.batch_status.numpy()
                still = []
                for k, a in enumerate(members):
                    a.steps += 1
                    st = self._parse_status(buf[k])
                    if st["done"]:
                        t_end = time.perf_counter() - t0
                        arr = a.req.arrival if realtime else 0.0
                        done[a.req.rid] = dict(latency=t_end - arr, finish=t_end, steps=a.steps, p=a.req.p,
                                               nel=a.req.nel, ndof=3 * a.req.nb_cp, **st)
                        if a.keep:
                            xs[a.req.rid] = a.x.cpu().numpy().copy()
                        self._release(a)
                        if progress:
                            progress(len(done), len(reqs))
                    else:
                        still.append(a)
                members = still
            torch.cuda.synchronize()
            makespan = time.perf_counter() - t0
        finally:
            stop[0] = True
            torch.cuda.synchronize()
            for a in members + joining:
                try:
                    self._release(a)
                except Exception:
                    pass
        lat = [d["latency"] for d in done.values()]
        rep = {"completed": len(done), "makespan_s": makespan,
               "solves_per_s": len(done) / makespan if makespan > 0 else None,
               "latency_s": latency_summary(lat), "failed": sum(1 for d in done.values() if d["failed"]),
               "at_fp64_floor": sum(1 for d in done.values() if d["floor"]),
               "peak_arena_bytes": self.alloc.peak, "budget_bytes": self.budget, "requests": done,
               "host_admission_s": t_admit, "step_wait_s": t_wait, "steps": nsteps}
        if keep_solutions:
            rep["x"] = xs
        return rep

    def close(self):
        if self.batch is not None:
            _lib.lib().yeti_batch_destroy(self.batch)
            self.batch = None
        self.arena_buf = None
