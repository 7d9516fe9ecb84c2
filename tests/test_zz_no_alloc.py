"""mgs_solve_noalloc, the solve entry point that allocates no device memory:
through the C-ABI it returns bitwise the field of the handle-based path
(mgs_create + mgs_set_density + mgs_solve + mgs_get_field) on two calls in a
row, the second on a smaller grid (the scratch array keeps the first call's
contents), for f(R) mixed, f(R) FP64 and nDGP; a warm start from a caller's
field; and it refuses a grid whose hierarchy exceeds its fixed scratch."""
import ctypes as C

import numpy as np
import pytest

import mgsolve as M
import oracle as O
from conftest import gpu_available

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    if not gpu_available():
        pytest.skip("no GPU")
    return torch


def _cfg(N, model, precision, box):
    return M.Config(N=N, model=model, fr_n=1, box=box, omega_m=0.3, omega_l=0.7, aexpn=1.0, fr0=-1e-5, h0rc=1.0,
                    precision=precision)


def _handle_solve(cfg, delta, opts, u0=None):
    with M.Solver(cfg) as S:
        S.set_density(delta)
        if u0 is not None:
            S.set_field(u0)
        r = M.SolveReport()
        rc = S.L.mgs_solve(S.h, C.byref(opts), C.byref(r))
        assert rc == 0, S.L.mgs_last_error(S.h).decode()
        return S.get_field(np.float64), r


def _noalloc_solve(torch, cfg, delta, opts, u0=None):
    L = M.load()
    N = cfg.N
    d_dev = torch.from_numpy(np.ascontiguousarray(delta, np.float64)).cuda()
    if u0 is None:
        u_dev = torch.full((N, N, N), float("nan"), dtype=torch.float64, device="cuda")
    else:
        u_dev = torch.from_numpy(np.ascontiguousarray(u0, np.float64)).cuda()
    torch.cuda.synchronize()
    p = cfg.c()
    r = M.SolveReport()
    print(f"[noalloc] N={N} model={cfg.model} precision={cfg.precision}", flush=True)
    rc = L.mgs_solve_noalloc(C.byref(p), C.byref(opts), C.c_void_p(d_dev.data_ptr()), M.F64,
                             C.c_void_p(u_dev.data_ptr()), M.F64, C.byref(r))
    assert rc == 0, L.mgs_last_error(None).decode()
    torch.cuda.synchronize()
    return u_dev.cpu().numpy(), r


CASES = [
    # (N, model, precision, box, sigma_g, seed)
    (64, M.MODEL_FR, M.PREC_MIXED, 128.0, 1.2, 11),
    (32, M.MODEL_FR, M.PREC_MIXED, 64.0, 1.0, 12),
    (48, M.MODEL_FR, M.PREC_FP64, 96.0, 1.0, 13),
    (32, M.MODEL_DGP, M.PREC_FP64, 64.0, 1.0, 14),
    (64, M.MODEL_DGP, M.PREC_MIXED, 128.0, 1.0, 15),
    (16, M.MODEL_DGP, M.PREC_MIXED, 32.0, 0.8, 16),
]


def test_noalloc_matches_handle_solve_on_consecutive_calls():
    torch = _torch()
    for N, model, prec, box, sg, seed in CASES:
        cfg = _cfg(N, model, prec, box)
        delta = O.lognormal_density(N, box, sg, seed=seed)
        opts = M.SolveOpts(1, 40, 1e-8, 1e-8, 1, 0, 0, 0)
        u_h, r_h = _handle_solve(cfg, delta, opts)
        u_n, r_n = _noalloc_solve(torch, cfg, delta, opts)
        assert r_n.converged == r_h.converged and r_n.cycles == r_h.cycles and r_n.outer == r_h.outer, (N, model)
        assert r_n.res0 == r_h.res0 and r_n.res_final == r_h.res_final, (N, model, r_n.res_final, r_h.res_final)
        assert np.array_equal(u_n, u_h), (N, model, prec, float(np.max(np.abs(u_n - u_h))))


def test_noalloc_warm_start_matches_handle():
    torch = _torch()
    N, box = 32, 64.0
    cfg = _cfg(N, M.MODEL_FR, M.PREC_MIXED, box)
    delta = O.lognormal_density(N, box, 1.0, seed=21)
    first = M.SolveOpts(1, 2, 1e-8, 1e-8, 1, 1, 0, 0)
    u0, _ = _handle_solve(cfg, delta, first)
    warm = M.SolveOpts(1, 40, 1e-8, 1e-8, 0, 0, 0, 0)
    u_h, r_h = _handle_solve(cfg, delta, warm, u0=u0)
    u_n, r_n = _noalloc_solve(torch, cfg, delta, warm, u0=u0)
    assert r_n.converged and r_n.outer == r_h.outer
    assert np.array_equal(u_n, u_h)


def test_noalloc_rejects_grids_beyond_its_scratch():
    torch = _torch()
    L = M.load()
    cfg = _cfg(128, M.MODEL_FR, M.PREC_MIXED, 256.0)
    p = cfg.c()
    opts = M.SolveOpts(1, 40, 1e-8, 1e-8, 1, 0, 0, 0)
    r = M.SolveReport()
    buf = torch.zeros(8, dtype=torch.float64, device="cuda")  # never touched: the call fails before any launch
    rc = L.mgs_solve_noalloc(C.byref(p), C.byref(opts), C.c_void_p(buf.data_ptr()), M.F64,
                             C.c_void_p(buf.data_ptr()), M.F64, C.byref(r))
    assert rc == -1
    assert "scratch" in L.mgs_last_error(None).decode()
