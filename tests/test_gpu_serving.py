"""GPU tests of the steppable solver sessions and the serving loop.

Bars: the served displacement against a direct solve of the oracle's K on the free dofs
within 1e-8 relative (mixed precision, rtol 1e-10) and the same for the FP64 sessions; graph
replays bit-identical to eager steps, also after eager work in between; a request's solution
bit-identical alone and inside a mix; arena footprint never above the budget."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from yeti_amd.geometry import make_box_patch

pytestmark = pytest.mark.gpu


def direct_u(args):
    K, r, c, F = O.oracle_sys_linmat((), kwargs=args)
    n = int(args["nb_dof_tot"])
    A = sp.coo_matrix((K, (r, c)), shape=(n, n)).tocsr()
    A = (A + A.T).tocsr()
    idf = np.asarray(args["ind_dof_free"]).astype(np.int64)[: int(args["nb_dof_free"])] - 1
    u = np.zeros(n)
    u[idf] = spla.spsolve(A[idf][:, idf].tocsc(), F[idf])
    return u


def small_requests():
    from yeti_amd.serving import Request
    rng = np.random.default_rng(3)
    out = []
    for i in range(12):
        nel = tuple(int(v) for v in rng.integers(2, 7, size=3))
        p = int(rng.choice([2, 3, 4]))
        out.append(Request(i, 0.0, nel, p, float(rng.uniform(1e3, 1e5)), float(rng.uniform(0.2, 0.4)),
                           -float(rng.uniform(0.1, 10)), int(rng.integers(1 << 31))))
    return out


@pytest.mark.parametrize("precision,precond", [("mixed", "fdm"), ("fp64", "fdm"), ("mixed", "jacobi")])
def test_served_solutions_match_direct_solve(gpu, precision, precond):
    from yeti_amd.serving import SolveServer
    print("[case] served solutions vs direct solve", precision, precond, flush=True)
    reqs = small_requests()[:6]
    srv = SolveServer(budget_bytes=1 << 30, precision=precision, nstreams=3, chunk=6, rtol=1e-10, precond=precond)
    rep = srv.run(reqs, keep_solutions=True, realtime=False)
    assert rep["completed"] == len(reqs) and rep["failed"] == 0
    for r in reqs:
        u = direct_u(r.model())
        err = np.linalg.norm(rep["x"][r.rid] - u) / np.linalg.norm(u)
        info = rep["requests"][r.rid]
        print("  req %d nel %s p %d: %d iterations, %d replacements, relres %.2e err %.2e" % (
            r.rid, r.nel, r.p, info["iters"], info["updates"], info["relres"], err))
        assert info["relres"] <= 1e-10
        assert err <= 1e-8
    assert rep["peak_arena_bytes"] <= srv.budget


def test_graph_replays_match_eager_steps(gpu):
    import torch
    from yeti_amd.serving import SolveServer
    print("[case] graph replay vs eager", flush=True)
    reqs = small_requests()[6:9]
    xs = []
    for use_graph in (False, True, True):
        srv = SolveServer(budget_bytes=1 << 30, precision="mixed", nstreams=2, chunk=5, use_graph=use_graph)
        rep = srv.run(reqs, keep_solutions=True, realtime=False)
        # eager work of the program between runs
        t = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        assert float((t * t).sum()) > 0
        xs.append(rep["x"])
    for r in reqs:
        assert np.array_equal(xs[0][r.rid], xs[1][r.rid])
        assert np.array_equal(xs[1][r.rid], xs[2][r.rid])


def test_result_identical_alone_and_in_mix(gpu):
    from yeti_amd.serving import SolveServer
    print("[case] alone vs mix", flush=True)
    reqs = small_requests()
    srv = SolveServer(budget_bytes=1 << 30, precision="mixed", nstreams=4, chunk=8)
    mix = srv.run(reqs, keep_solutions=True, realtime=False)
    for k in (0, 5, 11):
        alone = SolveServer(budget_bytes=1 << 30, precision="mixed", nstreams=1, chunk=8).run(
            [reqs[k]], keep_solutions=True, realtime=False)
        assert np.array_equal(alone["x"][k], mix["x"][k])


def test_budget_admission_limits_footprint(gpu):
    from yeti_amd.serving import SolveServer
    print("[case] budget admission", flush=True)
    reqs = small_requests()
    budget = 24 << 20
    srv = SolveServer(budget_bytes=budget, precision="mixed", nstreams=4, chunk=8)
    rep = srv.run(reqs, realtime=False)
    assert rep["completed"] == len(reqs)
    assert rep["peak_arena_bytes"] <= budget
    print("  peak arena %.1f MiB of %.1f MiB" % (rep["peak_arena_bytes"] / 2**20, budget / 2**20))


@pytest.mark.parametrize("precond", [0, 1])
def test_session_footprint_formula(gpu, precond):
    import ctypes as C
    import torch
    from yeti_amd import _lib
    from yeti_amd.matfree import MatrixFree
    from yeti_amd.serving import ir_session_bytes
    from yeti_amd.solver import free_mask_from_args
    print("[case] session footprint", precond, flush=True)
    bp = make_box_patch(nel=(5, 3, 2), p=3)
    mf = MatrixFree(bp.args)
    mask = free_mask_from_args(bp.args)
    dims = (5 + 3, 3 + 3, 2 + 3)
    for prec in (32, 64):
        arena = torch.empty(64 << 20, dtype=torch.uint8, device="cuda")
        h, used, counted = C.c_void_p(), C.c_size_t(0), C.c_size_t(0)
        _lib.check(_lib.lib().yeti_ir_create(mf._h, C.c_void_p(mask.data_ptr()), prec, precond, 1e-10, 0.1, 4, 1000,
                                             20, None, 0, None, C.byref(h), C.byref(counted)))
        _lib.check(_lib.lib().yeti_ir_create(mf._h, C.c_void_p(mask.data_ptr()), prec, precond, 1e-10, 0.1, 4, 1000,
                                             20, C.c_void_p(arena.data_ptr()), arena.numel(), None, C.byref(h),
                                             C.byref(used)))
        bound = ir_session_bytes(dims, prec, "fdm" if precond else "jacobi")
        assert used.value <= counted.value <= bound
        _lib.lib().yeti_ir_destroy(h)


# ---- batched serving: every active request advanced by the same launches -----------------


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batched_solutions_match_direct_solve(gpu, precision):
    """The batched server (shared launches walking device task lists) solves each request of a
    mix of degrees 2-4 to rtol 1e-10 on its FP64 residual, within 1e-8 of the direct solve."""
    from yeti_amd.serving import SolveServer
    print("[case] batched solutions vs direct solve", precision, flush=True)
    reqs = small_requests()
    srv = SolveServer(budget_bytes=1 << 30, precision=precision, chunk=6, rtol=1e-10, batched=True)
    rep = srv.run(reqs, keep_solutions=True, realtime=False)
    srv.close()
    assert rep["completed"] == len(reqs) and rep["failed"] == 0
    for r in reqs:
        u = direct_u(r.model())
        err = np.linalg.norm(rep["x"][r.rid] - u) / np.linalg.norm(u)
        info = rep["requests"][r.rid]
        print("  req %d nel %s p %d: %d iterations, %d replacements, relres %.2e err %.2e" % (
            r.rid, r.nel, r.p, info["iters"], info["updates"], info["relres"], err))
        assert info["relres"] <= 1e-10 or info["floor"]
        assert err <= 1e-8
    assert rep["peak_arena_bytes"] <= srv.budget


def test_batched_result_identical_alone_and_in_mix(gpu):
    """A request's batched solution is bit-identical alone and inside a batch with eleven
    others of other sizes and degrees, admitted together or arriving later (its reductions are
    chunked by its own size only), and equal with and without graph capture."""
    from yeti_amd.serving import SolveServer
    print("[case] batched alone vs mix", flush=True)
    reqs = small_requests()
    mix = SolveServer(budget_bytes=1 << 30, precision="mixed", chunk=8, batched=True).run(
        reqs, keep_solutions=True, realtime=False)
    # staggered admissions: members join while others are mid-solve
    staggered = [type(r)(r.rid, 0.002 * i, r.nel, r.p, r.E, r.nu, r.traction, r.seed) for i, r in enumerate(reqs)]
    for a, b in zip(staggered, reqs):
        a.args = b.args
    late = SolveServer(budget_bytes=1 << 30, precision="mixed", chunk=8, batched=True).run(
        staggered, keep_solutions=True, realtime=True)
    for k in (0, 5, 11):
        alone = SolveServer(budget_bytes=1 << 30, precision="mixed", chunk=8, batched=True).run(
            [reqs[k]], keep_solutions=True, realtime=False)
        eager = SolveServer(budget_bytes=1 << 30, precision="mixed", chunk=8, batched=True, use_graph=False).run(
            [reqs[k]], keep_solutions=True, realtime=False)
        assert np.array_equal(alone["x"][k], mix["x"][k])
        assert np.array_equal(alone["x"][k], late["x"][k])
        assert np.array_equal(alone["x"][k], eager["x"][k])


def test_batched_stream_requests_match_assembled_fp64(gpu):
    """Requests of the serving stream itself (seeded generator: 4-48 elements per direction,
    degrees 2-4, the bench's seed), the first eight whose patches have at most 24^3 elements,
    served by the batched server in mixed precision to rtol 1e-10: each displacement within 1e-8
    of an FP64 solve whose every product is the assembled CSR of K + K^T (device COO of the
    oracle-pinned assembly), solved to 1e-12 or its FP64 floor."""
    import torch
    from yeti_amd.assembly import Assembler
    from yeti_amd.matfree import MatrixFree
    from yeti_amd.serving import SolveServer, request_stream, solve_session
    from yeti_amd.solver import StiffnessCSR, free_mask_from_args
    print("[case] batched stream requests vs assembled FP64", flush=True)
    reqs = [r for r in request_stream(n=200, rate=40.0, seed=20261202) if np.prod(r.nel) <= 24 ** 3][:8]
    reqs = [type(r)(k, 0.0, r.nel, r.p, r.E, r.nu, r.traction, r.seed) for k, r in enumerate(reqs)]
    assert sorted({r.p for r in reqs}) == [2, 3, 4]
    srv = SolveServer(budget_bytes=8 << 30, precision="mixed", rtol=1e-10, batched=True)
    rep = srv.run(reqs, keep_solutions=True, realtime=False)
    srv.close()
    assert rep["completed"] == len(reqs) and rep["failed"] == 0
    for r in reqs:
        args = r.model()
        asm = Assembler(args)
        F = asm.run_rhs()
        K = torch.empty(asm.nb_data, dtype=torch.float64, device="cuda")
        asm.run_values(K)
        csr = StiffnessCSR(asm)
        vals = csr.values(K)
        del K
        mask = free_mask_from_args(args)
        mf = MatrixFree(args)
        u0, st0 = solve_session(mf, F, mask, precision=64, rtol=1e-12, csr=csr, csr_vals=vals, max_iters=50000)
        assert not st0["failed"]
        u0 = u0.cpu().numpy()
        err = np.linalg.norm(rep["x"][r.rid] - u0) / np.linalg.norm(u0)
        info = rep["requests"][r.rid]
        print("  req %d nel %s p %d: %d iterations, relres %.2e, |du|/|u| %.2e (FP64 %d its, relres %.2e)" % (
            r.rid, r.nel, r.p, info["iters"], info["relres"], err, st0["iters"], st0["relres"]), flush=True)
        assert err <= 1e-8
        mf.close()
        csr.close()
        asm.close()
