"""The continuous-batching solve server (mgs_server_* in include/mgsolve.h).

Each request's result must be the same alone, in a mix with other requests
of other sizes, parameters and models (f(R), nDGP), and from the
single-request solve (mgs_solve, mixed precision: f(R) requests stopping at
a relative tolerance as mode 1 with automatic inner-cycle counts, the others
as mode 0 or mode 1 with one inner cycle per outer iteration): the server
runs the same per-request arithmetic, tiling, reduction order and inner
counts, so the bar is bitwise equality.
Admission by HBM footprint is exercised with a budget that holds only one
request at a time.  Error paths are checked without a GPU."""
import time

import numpy as np
import pytest

import oracle as O
import mgsolve as M
from conftest import gpu_available


def cfg(N, fr0=-1e-5, frn=1, box=64.0, a=1.0):
    return M.Config(N=N, model=M.MODEL_FR, fr_n=frn, box=box, omega_m=0.3, omega_l=0.7, aexpn=a, fr0=fr0,
                    precision=M.PREC_MIXED)


# a mix of sizes (incl. a non-power-of-two grid), models n and parameters
MIX = [
    (32, -1e-5, 1, 64.0, 1.0, 1.0, 11),
    (16, -1e-4, 1, 32.0, 1.0, 0.8, 12),
    (48, -1e-6, 1, 96.0, 1.0, 1.3, 13),
    (32, -1e-5, 0, 64.0, 0.8, 1.0, 14),
    (64, -1e-5, 2, 128.0, 1.0, 1.1, 15),
    (24, -1e-5, 1, 48.0, 1.0, 1.2, 16),
    (32, -1e-6, 1, 64.0, 0.5, 1.5, 17),
]


def density(N, box, sigma, seed, f32):
    d = O.lognormal_density(N, box, sigma, seed=seed)
    return d.astype(np.float32) if f32 else d


def standalone(c, delta, tol):
    with M.Solver(c) as S:
        S.set_density(delta)
        rep = S.solve(mode=1, rel_tol=tol, max_cycles=60, init_guess=True, inner_cycles=0)
        return S.get_field(), rep


@pytest.mark.gpu
def test_results_equal_alone_in_mix_and_standalone():
    tol = 1e-8
    reqs = []
    for k, (N, fr0, frn, box, a, sigma, seed) in enumerate(MIX):
        reqs.append((cfg(N, fr0, frn, box, a), density(N, box, sigma, seed, f32=(k % 2 == 0))))
    print(f"[serving] {len(reqs)} requests, mixed then alone", flush=True)
    with M.Server(max_active=4) as srv:
        tickets = [srv.submit(c, d, rel_tol=tol) for c, d in reqs]
        mixed = [srv.wait(t, c.N) for t, (c, _) in zip(tickets, reqs)]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["max_batch"] >= 2, st
    for (c, d), (u_mix, r_mix) in zip(reqs, mixed):
        print(f"[serving] N={c.N} n={c.fr_n} fr0={c.fr0}", flush=True)
        assert r_mix["status"] == 0 and r_mix["converged"], r_mix
        assert r_mix["res_final"] < tol * r_mix["res0"]
        with M.Server(max_active=1) as srv:
            u_alone, r_alone = srv.wait(srv.submit(c, d, rel_tol=tol), c.N)
        assert np.array_equal(u_alone, u_mix)
        assert r_alone["outer"] == r_mix["outer"] and r_alone["res_final"] == r_mix["res_final"]
        u_ref, rep = standalone(c, d, tol)
        assert np.array_equal(u_ref, u_mix)
        assert rep["outer"] == r_mix["outer"] and rep["res_final"] == r_mix["res_final"]


@pytest.mark.gpu
@pytest.mark.parametrize("use_graphs", [True, False])
def test_admission_by_footprint_and_graph_replay(use_graphs):
    """A budget that holds one 32^3 hierarchy: requests queue and are
    admitted one after another; results do not depend on graph replay.
    Results wait in HBM until claimed, so admitting the next request spills
    the unclaimed one to the host.  Waiting for the last request first
    makes the path of each certain: request 2 is claimed from HBM, requests
    0 and 1 (spilled to admit 1 and 2) are read from their host copies; as
    FP64 or FP32 and scaled, every path gives the field the unconstrained
    server copies out of HBM."""
    N = 32
    c = cfg(N)
    ds = [density(N, 64.0, 1.0, 40 + i, f32=True) for i in range(3)]
    print(f"[serving] admission, graphs={use_graphs}", flush=True)
    with M.Server(max_active=8, use_graphs=True) as big:
        ref = [big.wait(big.submit(c, d), N)[0] for d in ds]
    budget = 40 * N ** 3  # one hierarchy (about 35 bytes per cell), not two
    reads = [(np.float64, 1.0), (np.float32, 1.0), (np.float32, 0.25)]
    with M.Server(max_active=8, hbm_budget=budget, use_graphs=use_graphs) as srv:
        ts = [srv.submit(c, d) for d in ds]
        while not srv.test(ts[2]):  # non-blocking completion check
            time.sleep(0.001)
        last = srv.wait(ts[2], N, *reads[2])
        out = [srv.wait(t, N, dt, sc) for t, (dt, sc) in zip(ts[:2], reads[:2])] + [last]
        st = srv.stats()
        with pytest.raises(M.MgsError, match="unknown ticket"):
            srv.test(ts[0])  # claimed tickets are gone
    assert st["max_batch"] == 1 and st["spilled"] == 2, st
    for (u, r), u_ref, (dt, sc) in zip(out, ref, reads):
        assert r["status"] == 0 and r["converged"]
        assert u.dtype == dt
        assert np.array_equal(u, (u_ref * sc).astype(dt))
    # requests queued behind the first waited for admission
    assert out[2][1]["queued_ms"] >= out[0][1]["queued_ms"]


def dgp_cfg(N, branch=1, box=64.0, a=1.0, h0rc=1.0):
    return M.Config(N=N, model=M.MODEL_DGP, branch=branch, box=box, omega_m=0.3, omega_l=0.7, aexpn=a, h0rc=h0rc,
                    precision=M.PREC_MIXED)


# nDGP requests (power-of-two grids, both branches) mixed with f(R) ones
DGP_MIX = [
    (32, 1, 64.0, 1.0, 1.0, 1.0, 21),
    (16, -1, 32.0, 1.0, 1.0, 0.8, 22),
    (64, 1, 128.0, 0.8, 2.0, 1.1, 23),
    (32, -1, 64.0, 1.0, 0.5, 1.2, 24),
    (128, 1, 256.0, 1.0, 1.0, 1.0, 25),  # the standalone solve's z-marching passes
]
DGP_EXIT = dict(exit_rule=M.EXIT_REFERENCE, res_conv=1e-30, max_outer=40)


@pytest.mark.gpu
def test_dgp_results_equal_alone_in_mix_and_standalone():
    """nDGP requests stop by the reference's rule (the DGP residual of a
    periodic box levels off, so the stagnation exits end the solve); in a
    batch with f(R) requests each result is bitwise that of the request
    alone and of mgs_solve (mixed, mode 0, one inner cycle; the f(R)
    requests of the batch as mode 1 with automatic inner counts)."""
    tol = 1e-8
    reqs = []
    for k, (N, branch, box, a, h0rc, sigma, seed) in enumerate(DGP_MIX):
        reqs.append(("dgp", dgp_cfg(N, branch, box, a, h0rc), density(N, box, sigma, seed, f32=(k % 2 == 1))))
    for k, (N, fr0, frn, box, a, sigma, seed) in enumerate(MIX[:2]):
        reqs.append(("fr", cfg(N, fr0, frn, box, a), density(N, box, sigma, seed, f32=True)))

    def submit(srv, kind, c, d):
        return srv.submit(c, d, **DGP_EXIT) if kind == "dgp" else srv.submit(c, d, rel_tol=tol)

    print(f"[serving] {len(reqs)} nDGP + f(R) requests, mixed", flush=True)
    with M.Server(max_active=6) as srv:
        tickets = [submit(srv, k, c, d) for k, c, d in reqs]
        mixed = [srv.wait(t, c.N) for t, (_, c, _) in zip(tickets, reqs)]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["max_batch"] >= 4, st
    for (kind, c, d), (u_mix, r_mix) in zip(reqs, mixed):
        print(f"[serving] {kind} N={c.N} outer={r_mix['outer']} res {r_mix['res0']:.3e} -> "
              f"{r_mix['res_final']:.3e}", flush=True)
        assert r_mix["status"] == 0, r_mix
        assert r_mix["res_final"] < 0.05 * r_mix["res0"], r_mix
        with M.Server(max_active=1) as srv:
            u_alone, r_alone = srv.wait(submit(srv, kind, c, d), c.N)
        assert np.array_equal(u_alone, u_mix)
        assert r_alone["outer"] == r_mix["outer"] and r_alone["res_final"] == r_mix["res_final"]
        with M.Solver(c) as S:
            S.set_density(d)
            if kind == "dgp":
                rep = S.solve(mode=0, res_conv=1e-30, max_cycles=40, inner_cycles=1)
            else:
                rep = S.solve(mode=1, rel_tol=tol, max_cycles=60, inner_cycles=0)
            u_ref = S.get_field()
        assert np.array_equal(u_ref, u_mix)
        assert rep["outer"] == r_mix["outer"] and rep["res_final"] == r_mix["res_final"]
        assert rep["converged"] == r_mix["converged"]


@pytest.mark.gpu
def test_spill_and_claim_with_concurrent_submit_and_wait():
    """One thread submits while another waits in shuffled order, under a
    budget of about three 48^3 hierarchies: finished results wait in HBM, so
    admission keeps spilling unclaimed ones to the host while waiters claim
    others from HBM.  Waiting for the last ticket first makes spills certain
    (every earlier request had to be admitted before it).  Every field, read
    as FP64 or FP32, equals the one an unconstrained server returns."""
    import threading
    import time

    sizes = [48, 32, 48, 48, 32, 48, 48, 48]
    base = [(cfg(N, -1e-5, 1, 2.0 * N), density(N, 2.0 * N, 1.0 + 0.05 * q, 60 + q, f32=(q % 2 == 0)))
            for q, N in enumerate(sizes)]
    reqs = [base[q % len(base)] for q in range(3 * len(base))]
    print(f"[serving] concurrent submit/wait, {len(reqs)} requests", flush=True)
    with M.Server(max_active=8) as big:
        ref = [big.wait(big.submit(c, d), c.N)[0] for c, d in base]
    budget = 3 * 40 * 48 ** 3
    rng = np.random.default_rng(3)
    tickets = [None] * len(reqs)
    with M.Server(max_active=4, hbm_budget=budget) as srv:

        def producer():
            for q, (c, d) in enumerate(reqs):
                tickets[q] = srv.submit(c, d)
                time.sleep(float(rng.uniform(0.0, 0.002)))

        th = threading.Thread(target=producer)
        th.start()
        order = [len(reqs) - 1] + [int(q) for q in rng.permutation(len(reqs) - 1)]
        out = {}
        for q in order:
            while tickets[q] is None:
                time.sleep(0.0005)
            dt = np.float64 if q % 2 else np.float32
            out[q] = srv.wait(tickets[q], reqs[q][0].N, dt)
        th.join()
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["spilled"] > 0, st
    for q in range(len(reqs)):
        u, r = out[q]
        assert r["status"] == 0 and r["converged"], (q, r)
        assert np.array_equal(u, ref[q % len(base)].astype(u.dtype)), q


@pytest.mark.gpu
def test_config4_stream_meets_the_oracle_operator_alone_equals_mix():
    """The serving stream's workload: f(R) n=1 requests with N drawn from
    {64, 128, 192, 256}, |fR0| from {1e-4, 1e-5, 1e-6}, sigma_g from
    U[0.8, 1.6], seeded lognormal densities, arriving while others run.
    Every served field meets the oracle's FP64 operator at relative residual
    1e-8 (+5 % for the norm's summation order); four of them (one per size)
    are bitwise the same served alone; two carry the FP64 comparison solve
    (verify_fp64), whose field they match to 1e-6."""
    rng = np.random.default_rng(404)
    sizes = [64, 128, 192, 256] * 6
    rng.shuffle(sizes)
    reqs = []
    for i, N in enumerate(sizes):
        fr0 = -float(rng.choice([1e-4, 1e-5, 1e-6]))
        sg = float(rng.uniform(0.8, 1.6))
        c = cfg(N, fr0=fr0, box=2.0 * N)
        with M.Solver(M.Config(N=N, model=M.MODEL_FR, box=2.0 * N, precision=M.PREC_MIXED)) as G:
            G.make_lognormal_density(2000 + i, sg)
            d = G.get_density(np.float64)
        reqs.append((c, d))
    print(f"[serving] config-4 stream: {len(reqs)} requests", flush=True)
    tol = 1e-8
    verify = {3, 17}
    with M.Server(max_active=16) as srv:
        tickets = []
        for i, (c, d) in enumerate(reqs):
            tickets.append(srv.submit(c, d, rel_tol=tol, verify_fp64=(i in verify)))
            time.sleep(float(rng.exponential(0.002)))  # Poisson arrivals
        out = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["max_batch"] >= 4, st
    for i, ((c, d), (u, r)) in enumerate(zip(reqs, out)):
        assert r["status"] == 0 and r["converged"], (i, r)
        op = O.Params(N=c.N, model=1, fr_n=1, branch=1, box=c.box, om=0.3, oml=0.7, a=1.0, fr0=c.fr0, h0rc=1.0)
        res = O.fr_residual(op, 0, u, d, None, None, "f64")
        r0 = 0.3 * float(np.sqrt(np.mean(d * d)))  # residual of u = u_bg: (Om/a) delta
        assert abs(r["res0"] / r0 - 1.0) < 1e-10, (i, r["res0"], r0)
        assert res < 1.05 * tol * r0, (i, c.N, res / r0)
        if i in verify:
            assert r["verified"] == 1 and r["fp64_rel_linf"] < 1e-6, (i, r)
            assert r["fp64_res_final"] < tol * r["res0"], (i, r)
        else:
            assert r["verified"] == 0
    seen = {}
    for i, (c, _) in enumerate(reqs):
        seen.setdefault(c.N, i)
    with M.Server(max_active=16) as alone:
        for N, i in sorted(seen.items()):
            c, d = reqs[i]
            u_alone, _ = alone.wait(alone.submit(c, d, rel_tol=tol))
            assert np.array_equal(u_alone, out[i][0]), (N, i)


@pytest.mark.gpu
def test_wait_misuse_is_an_error_not_a_fault():
    """A second wait on one ticket, an output array smaller than the field,
    and the wrong N or dtype are refused; destroying a server with a waiter
    blocked on an unfinished request wakes it with an error."""
    import ctypes as C
    import threading

    d = density(32, 64.0, 1.0, 3, f32=True)
    with M.Server(max_active=2) as srv:
        t = srv.submit(cfg(32), d)
        with pytest.raises(M.MgsError, match="submitted with N"):
            srv.wait(t, 16)
        with pytest.raises(M.MgsError, match="dtype"):
            srv.wait(t, 32, np.int32)
        small = np.empty(16 ** 3, np.float64)
        res = M.RequestResult()
        rc = srv.L.mgs_server_wait(srv.h, t, small.ctypes.data_as(C.c_void_p), small.size, M.F64, 1.0,
                                   C.byref(res))
        assert rc == -1 and "elements" in srv.L.mgs_server_last_error(srv.h).decode()
        u, r = srv.wait(t)
        assert r["status"] == 0 and u.shape == (32, 32, 32)
        with pytest.raises(M.MgsError, match="unknown ticket"):
            srv.wait(t, 32)
        # two waiters on one ticket: the second is refused while the first blocks
        t2 = srv.submit(cfg(64, box=128.0), density(64, 128.0, 1.2, 4, f32=True))
        first = {}
        th = threading.Thread(target=lambda: first.setdefault("r", srv.wait(t2)))
        th.start()
        time.sleep(0.01)
        rc2 = srv.L.mgs_server_wait(srv.h, t2, None, 0, M.F64, 1.0, None)
        th.join()
        assert rc2 in (-4, -1), rc2  # already waited for (-4) or, if the first finished, unknown (-1)
        assert first["r"][1]["status"] == 0
    # a full queue refuses a submission (its staging goes back to the pool
    # without re-taking the server's lock) and the accepted requests finish
    with M.Server(max_active=1, queue_capacity=1) as srv:
        dd = density(128, 256.0, 1.5, 6, f32=True)
        accepted, busy = [], 0
        for _ in range(4):
            try:
                accepted.append(srv.submit(cfg(128, box=256.0), dd, rel_tol=1e-10))
            except M.MgsError as e:
                assert "queue full" in str(e)
                busy += 1
        assert busy >= 1 and accepted
        for t in accepted:
            assert srv.wait(t)[1]["status"] == 0
        assert srv.stats()["rejected"] == busy
    # destroy with a blocked waiter
    srv = M.Server(max_active=1)
    t3 = srv.submit(cfg(128, box=256.0), density(128, 256.0, 1.5, 5, f32=True), rel_tol=1e-12, max_outer=200)
    box = {}
    handle = srv.h
    entering = threading.Event()

    def waiter():
        r = M.RequestResult()
        entering.set()
        box["rc"] = srv.L.mgs_server_wait(handle, t3, None, 0, M.F64, 1.0, C.byref(r))

    th = threading.Thread(target=waiter)
    th.start()
    entering.wait()
    time.sleep(0.05)  # the waiter is inside mgs_server_wait well before this
    srv.close()
    th.join(timeout=60)
    assert not th.is_alive()
    assert box["rc"] in (0, -4), box


@pytest.mark.gpu
def test_server_rejects_bad_requests():
    with M.Server(max_active=2) as srv:
        d = np.zeros((16,) * 3, np.float32)
        with pytest.raises(M.MgsError, match="MGS_PREC_MIXED"):
            srv.submit(M.Config(N=16, precision=M.PREC_FP64), d)
        with pytest.raises(M.MgsError, match="ngrid"):
            srv.submit(cfg(8), np.zeros((8,) * 3, np.float32))
        with pytest.raises(M.MgsError, match="power-of-two"):
            srv.submit(dgp_cfg(24), np.zeros((24,) * 3, np.float32), **DGP_EXIT)
        with pytest.raises(M.MgsError, match="exit_rule"):
            srv.submit(dgp_cfg(16), d, exit_rule=7)
        with pytest.raises(M.MgsError, match="res_conv"):
            srv.submit(dgp_cfg(16), d, exit_rule=M.EXIT_REFERENCE, res_conv=-1.0)
        bad = d.copy()
        bad[1, 2, 3] = np.inf
        with pytest.raises(M.MgsError, match=r"finite \(cell 291\)"):  # (1 * 16 + 2) * 16 + 3
            srv.submit(cfg(16), bad)
        # a field copied by several threads: the bad value in the last slice
        big = np.zeros((192,) * 3, np.float32)
        big[191, 190, 7] = np.nan
        cell = (191 * 192 + 190) * 192 + 7
        with pytest.raises(M.MgsError, match=rf"finite \(cell {cell}\)"):
            srv.submit(cfg(192, box=384.0), big)
        assert srv.stats()["submitted"] == 0
        with pytest.raises(M.MgsError, match="budget"):
            with M.Server(max_active=1, hbm_budget=1024) as tiny:
                tiny.submit(cfg(16), d)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_server_without_gpu_is_an_error():
    with pytest.raises(M.MgsError, match="no CPU fallback"):
        M.Server()
