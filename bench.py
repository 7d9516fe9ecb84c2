"""Benchmark: f(R) solve time to relative residual 1e-8 at 512^3, mixed vs FP64.

Workload (default): Hu-Sawicki f(R) n=1, |fR0| = 1e-6, Om = 0.3, OmL = 0.7,
a = 1, box 512 Mpc/h, 512^3 grid; density delta = exp(g - sigma^2/2) - 1 with g
a seeded Gaussian field, P_g(k) ~ k^-1.5 exp(-k^2/(2 k_c^2)), k_c = k_Nyq/2,
sigma_g = 1.5, generated on the GPU.  One step = one solve from u = u_bg to
RMS(r)/RMS(r0) < 1e-8 with the chosen (mixed-precision) form; the plain FP64
solve of the same density is timed beside it.  Inputs are resident in HBM
before the timed region; the solve includes no host transfer.

Prints one JSON line (the last line of stdout).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mg-glam-durham_amd"))

import mgsolve as M  # noqa: E402

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (N, box, fr0, sigma_g, seed)
    "fr512": (512, 512.0, -1.0e-6, 1.5, 1002),
    "fr128": (128, 256.0, -1.0e-5, 1.0, 1001),
    "fr1024": (1024, 1024.0, -1.0e-5, 1.2, 1003),
    "fr256": (256, 256.0, -1.0e-6, 1.5, 1002),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_solver(N, box, fr0, precision):
    cfg = M.Config(N=N, model=M.MODEL_FR, fr_n=1, box=box, omega_m=0.3, omega_l=0.7, aexpn=1.0, fr0=fr0,
                   precision=precision)
    return M.Solver(cfg)


PROF_STEPS = 2  # profiled solves for the per-kernel classes and the roofline


def timed_solves(S, steps, warmup, profile, inner):
    for _ in range(warmup):
        S.solve(mode=1, rel_tol=1e-8, max_cycles=100, init_guess=True, inner_cycles=inner)
    S.synchronize()
    if profile:
        S.profile(True)
    reps = []
    walls = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        r = S.solve(mode=1, rel_tol=1e-8, max_cycles=100, init_guess=True, inner_cycles=inner)
        walls.append(time.perf_counter() - ts)
        reps.append(r)
    S.synchronize()
    total = time.perf_counter() - t0
    prof = None
    if profile:
        prof = {M.KC_NAMES[k]: S.profile_get(k) for k in range(6)}
        S.profile(False)
    return reps, walls, total, prof


def host_cpu_info():
    """CPU model, sockets and the logical CPUs this process may run on"""
    model, sockets = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    return {"model": model, "sockets": max(len(sockets), 1), "logical_cpus_visible": visible}


def cpu_fp64_solve_baseline(delta64, N, box, fr0, budget_s=60.0, max_repeats=3):
    """The oracle's FP64 solve (a C restatement of the reference's V(2,2)
    cycle with OpenMP loops, refined closed-form roots) of the same density
    to relative residual 1e-8 on the host cores, OMP_PROC_BIND=close,
    OMP_PLACES=cores; median over up to max_repeats solves within budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
        O.lib()
    except Exception as e:
        log(f"cpu fp64 baseline unavailable: {e}")
        return None
    p = O.Params(N=N, model=1, fr_n=1, branch=1, box=box, om=0.3, oml=0.7, a=1.0, fr0=fr0, h0rc=1.0, polish=1)
    times, rep = [], None
    t_all = time.perf_counter()
    while len(times) < max_repeats:
        t0 = time.perf_counter()
        _, rep = O.solve(p, delta64, mode=1, rel_tol=1e-8, max_cycles=60, prec="f64")
        times.append(time.perf_counter() - t0)
        log(f"[bench] cpu fp64 solve {len(times)}: {times[-1]:.2f} s, {rep['cycles']} cycles, "
            f"rel. residual {rep['res_final'] / rep['res0']:.2e}")
        if time.perf_counter() - t_all + times[-1] > budget_s:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    info = host_cpu_info()
    return {
        "value": float(np.median(times)),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": (f"whole FP64 solve of the same {N}^3 density to rel. residual 1e-8 by the oracle (C restatement "
                   f"of the reference's V(2,2) cycle, OpenMP, {threads} threads, OMP_PROC_BIND=close, "
                   f"OMP_PLACES=cores): {rep['cycles']} cycles, rel. residual "
                   f"{rep['res_final'] / rep['res0']:.2e}; median of {len(times)}"),
        "solves_s": times,
        "cpu": info,
    }


def cpu_reference_baseline(delta32, N, box, fr0, gpu_cycles, ncycles):
    """The reference's own V_cycle (flang -O2 -fopenmp, FP32 storage) on the
    host cores: a bounded sample of `ncycles` V-cycles on the same density,
    scaled to the cycle count of the FP64 solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
        R = O.RefLib(omp=True)
    except Exception as e:  # reference build not present
        log(f"cpu baseline unavailable: {e}")
        return None
    p = O.Params(N=N, model=1, fr_n=1, box=box, om=O.f32(0.3), oml=O.f32(0.7), a=1.0, fr0=fr0)
    R.init(p, iter_max=ncycles)
    R.put(1, delta32)
    R.vcycles(1)  # warm-up cycle (first touch of FI3)
    t0 = time.perf_counter()
    hist = R.vcycles(ncycles)
    dt = (time.perf_counter() - t0) / ncycles
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": dt * gpu_cycles,
        "unit": "s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"{ncycles} V-cycles of the reference's V_cycle at {N}^3 on the same density "
                   f"({dt:.3f} s/cycle, {threads} OpenMP threads), x {gpu_cycles} cycles of the FP64 solve; "
                   f"the reference's FP32 storage floors near rel. residual 1e-5 "
                   f"(its residual after the sample: {hist[-1]:.3e})"),
    }


def dgp_compare(N, box, sigma, seed, steps):
    """nDGP normal branch, H0rc = 1, a = 1: the plain FP64 solve and the mixed
    solve of the same seeded lognormal density, each run to the reference's
    stagnation exit (the residual of a periodic DGP problem levels off at the
    mean of its nonlinear source, so a relative 1e-8 target does not apply);
    times, cycles, final residuals and the field / fifth-force P(k) agreement."""
    out = {}
    fields = {}
    delta = None
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        c = M.Config(N=N, model=M.MODEL_DGP, branch=1, box=box, omega_m=0.3, omega_l=0.7, aexpn=1.0, h0rc=1.0,
                     precision=prec)
        with M.Solver(c) as S:
            if delta is None:
                S.make_lognormal_density(seed, sigma)
                delta = S.get_density(np.float64)
            else:
                S.set_density(delta)
            S.solve(mode=0, res_conv=1e-30, max_cycles=60)
            S.synchronize()
            reps = [S.solve(mode=0, res_conv=1e-30, max_cycles=60) for _ in range(steps)]
            fields[prec] = S.get_field()
            _, pk, _ = S.power_spectrum(2)
        r = reps[-1]
        key = "fp64" if prec == M.PREC_FP64 else "mixed"
        out[key] = {"ms_median": float(np.median([x["seconds"] * 1e3 for x in reps])), "cycles": r["cycles"],
                    "outer": r["outer"], "res0": r["res0"], "res_final": r["res_final"]}
        out[key + "_pk"] = pk
    # phi is defined up to a constant (L involves only its differences)
    u64 = fields[M.PREC_FP64] - fields[M.PREC_FP64].mean()
    um = fields[M.PREC_MIXED] - fields[M.PREC_MIXED].mean()
    d = np.abs(um - u64)
    pk64, pkm = out.pop("fp64_pk"), out.pop("mixed_pk")
    out["mixed_vs_fp64"] = {
        "u_rel_linf_mean_removed": float(d.max() / np.abs(u64).max()),
        "u_rel_l2_mean_removed": float(np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(u64 ** 2))),
        "fifth_force_pk_max_abs_ratio_minus_1": float(np.max(np.abs(pkm[pk64 > 0] / pk64[pk64 > 0] - 1.0))),
    }
    out["speedup_mixed_vs_fp64"] = out["fp64"]["ms_median"] / out["mixed"]["ms_median"]
    out["workload"] = (f"nDGP normal branch, H0rc=1, Om=0.3, a=1, {N}^3, box {box:g} Mpc/h, lognormal sigma_g={sigma} "
                       f"seed={seed}, both solved to the reference's stagnation exit")
    log(f"[dgp] fp64 {out['fp64']['ms_median']:.1f} ms ({out['fp64']['cycles']} cycles), mixed "
        f"{out['mixed']['ms_median']:.1f} ms ({out['mixed']['outer']} outer), u rel L_inf "
        f"{out['mixed_vs_fp64']['u_rel_linf_mean_removed']:.2e} (mean removed)")
    return out


def pm_step(N, seed, steps):
    """One PM step around the solve, as the reference's time step runs it
    (PMP2main.f90:39-58): CIC deposit of N^3 particles, the mixed f(R) solve,
    the Poisson FFT, the fifth-force potential and the push.  Particles start
    on a lattice displaced by up to half a cell (seeded); device times per
    operation from the library's hipEvent profiler, algorithmic bytes: deposit
    12 B per particle + 8 B per cell, potential 16 B per cell through the
    FFTs' traffic not counted, fifth force 16 B per cell, push 48 B per
    particle (the potential gathers hit the caches)."""
    rng = np.random.default_rng(seed)
    g = np.arange(1, N + 1, dtype=np.float32)
    pos = []
    for ax in range(3):  # x varies fastest over the particle index, like the grid
        shape = [1, 1, 1]
        shape[2 - ax] = N
        base = np.broadcast_to(g.reshape(shape), (N, N, N)).reshape(-1)
        p = base + rng.uniform(-0.5, 0.5, N ** 3).astype(np.float32)
        p = np.where(p < 1.0, p + N, p)
        p = np.where(p >= N + 1.0, p - N, p).astype(np.float32)
        pos.append(p)
        del base
    vel = [rng.normal(0.0, 0.05, N ** 3).astype(np.float32) for _ in range(3)]
    c = M.Config(N=N, model=M.MODEL_FR, fr_n=1, box=float(N), omega_m=0.3, omega_l=0.7, aexpn=1.0, fr0=-1e-6,
                 precision=M.PREC_MIXED)
    pc, xc = -0.5 * 0.02, 0.02
    out = {}
    with M.Solver(c) as S:
        S.set_particles(*pos, *vel)
        del pos, vel
        times = {"deposit": [], "solve": [], "potential": [], "fifth_force": [], "move": []}
        for it in range(steps + 1):
            S.profile(True)
            last = 0.0
            per = {}
            for name, fn in (("deposit", S.cic_deposit),
                             ("solve", lambda: S.solve(mode=1, rel_tol=1e-8, max_cycles=60)),
                             ("potential", lambda: S.potential(1.5 * 0.3)),
                             ("fifth_force", S.potential_add_fifth_force),
                             ("move", lambda: S.move(pc, xc))):
                t0 = time.perf_counter()
                r = fn()
                S.synchronize()
                wall = (time.perf_counter() - t0) * 1e3
                if name == "solve":
                    per[name] = r["seconds"] * 1e3
                else:
                    tot = S.profile_get(6)[1]
                    per[name] = tot - last
                    last = tot
                per[name + "_wall"] = wall
            S.profile(False)
            if it > 0:
                for k in times:
                    times[k].append(per[k])
        n3 = float(N) ** 3
        med = {k: float(np.median(v)) for k, v in times.items()}
        alg = {"deposit": 12.0 * n3 + 8.0 * n3, "fifth_force": 16.0 * n3, "move": 48.0 * n3}
        out = {
            "workload": f"{N}^3 grid, {N}^3 particles (lattice + U(-0.5, 0.5) cell displacements), f(R) n=1 "
                        f"|fR0|=1e-6, mixed solve to 1e-8",
            "ms": med,
            "alg_GBs": {k: alg[k] / (med[k] * 1e-3) / 1e9 for k in alg},
            "particles_per_s": {"deposit": n3 / (med["deposit"] * 1e-3), "move": n3 / (med["move"] * 1e-3)},
            "step_ms": sum(med.values()),
        }
    log(f"[pm] {N}^3: " + ", ".join(f"{k} {v:.2f} ms" for k, v in med.items()))
    return out


def lognormal_set(cfgs, sigmas, seed0):
    """Seeded lognormal densities (request i: seed seed0 + i), generated on the
    GPU and kept on the host as float32."""
    dens = []
    gens = {}
    for i, c in enumerate(cfgs):
        g = gens.get((c.N, c.box))
        if g is None:
            g = gens[(c.N, c.box)] = M.Solver(M.Config(N=c.N, box=c.box, precision=M.PREC_FP64))
        g.make_lognormal_density(seed0 + i, float(sigmas[i]))
        dens.append(g.get_density(np.float32))
    for g in gens.values():
        g.close()
    return dens


def serve_run(tag, cfgs, kws, dens, max_active, rng):
    """All requests offered at once (saturated throughput), then Poisson
    arrivals at 0.8x that rate (latency); in both a thread submits while the
    client claims the fields as they finish.  Four requests are re-solved
    alone and compared with their in-mix results.  kws[i]: submit arguments
    of request i (its exit rule); success: status 0 and, for the rel_tol
    rule, converged."""
    import threading

    n_req = len(cfgs)

    def ok(o, kw):
        return o["status"] == 0 and (o["converged"] or kw.get("exit_rule") == M.EXIT_REFERENCE)

    def stream(srv, gaps, keep):
        """One thread submits (after gaps[i] seconds for request i; none: at
        once) while this one claims the fields as they finish, into one
        reused output array per grid size; returns the wall time, the submit
        and delivery times, the result dicts and the kept fields."""
        tickets = [None] * n_req
        t_sub = [0.0] * n_req
        t_got = [0.0] * n_req
        res = [None] * n_req
        kept = {}

        def producer():
            start = time.perf_counter()
            at = 0.0
            for i in range(n_req):
                if gaps is not None:
                    at += gaps[i]
                    dt = start + at - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                t_sub[i] = time.perf_counter()
                tickets[i] = srv.submit(cfgs[i], dens[i], **kws[i])

        outbuf = {N: np.empty((N, N, N), np.float32) for N in sorted({c.N for c in cfgs})}
        th = threading.Thread(target=producer)
        tp = time.perf_counter()
        th.start()
        pending = list(range(n_req))
        while pending:
            ready = [i for i in pending if tickets[i] is not None and srv.test(tickets[i])]
            if not ready:
                time.sleep(0.0002)
                continue
            for i in ready:
                u, o = srv.wait(tickets[i], cfgs[i].N, np.float32, out=outbuf[cfgs[i].N])
                t_got[i] = time.perf_counter()
                res[i] = o
                if i in keep:
                    kept[i] = u.copy()
            done = set(ready)
            pending = [i for i in pending if i not in done]
        th.join()
        return time.perf_counter() - tp, t_sub, t_got, res, kept

    # saturated: every request offered at once
    with M.Server(max_active=max_active) as srv:
        sat_s, _, _, outs, _ = stream(srv, None, set())
        sat = srv.stats()
    rate = 0.8 * n_req / sat_s
    log(f"[{tag}] saturated: {n_req / sat_s:.1f} solves/s, p50 {sat['p50_ms']:.1f} ms, batch <= {sat['max_batch']}")
    bad = [o for o, kw in zip(outs, kws) if not ok(o, kw)]
    # Poisson arrivals at 0.8 x the saturated rate
    gaps = rng.exponential(1.0 / rate, n_req)
    keep = {0, 1, n_req // 2, n_req - 1}
    with M.Server(max_active=max_active) as srv:
        pois_s, t_sub, t_got, outs_p, fields = stream(srv, gaps, keep)
        st = srv.stats()
    bad += [o for o, kw in zip(outs_p, kws) if not ok(o, kw)]
    # delivered: submit call -> the field in the caller's array (wait() copies
    # it out of HBM; the client claims results as they finish)
    dl = 1e3 * (np.array(t_got) - np.array(t_sub))
    log(f"[{tag}] Poisson {rate:.1f}/s: ready in HBM p50 {st['p50_ms']:.1f} ms p99 {st['p99_ms']:.1f} ms; "
        f"delivered p50 {np.percentile(dl, 50):.1f} ms p99 {np.percentile(dl, 99):.1f} ms; "
        f"{st['solves_per_s']:.1f} solves/s, batch <= {st['max_batch']}")
    diff = 0.0
    with M.Server(max_active=1) as srv:
        for i in sorted(keep):
            u, _ = srv.wait(srv.submit(cfgs[i], dens[i], **kws[i]), cfgs[i].N, np.float32)
            diff = max(diff, float(np.max(np.abs(u - fields[i]))))
    return {
        "requests": n_req,
        "max_active": max_active,
        "saturated_solves_per_s": n_req / sat_s,
        "arrival_rate_per_s": rate,
        "p50_ms": st["p50_ms"],
        "p99_ms": st["p99_ms"],
        "mean_ms": st["mean_ms"],
        "latency_is": "submit -> result ready in HBM (server clock)",
        "delivered_p50_ms": float(np.percentile(dl, 50)),
        "delivered_p99_ms": float(np.percentile(dl, 99)),
        "solves_per_s": st["solves_per_s"],
        "wall_s": pois_s,
        "max_batch": st["max_batch"],
        "failed_or_unconverged": len(bad),
        "alone_vs_mix_max_abs_diff": diff,
        # the saturated leg's worker time (admission runs beside steps while
        # requests are active, so it overlaps step_ms)
        "saturated_worker_ms": {k: sat[k] for k in ("step_ms", "admit_ms", "finish_ms")},
        "saturated_steps": sat["steps"],
    }


def serving_stream(n_req, seed, max_active):
    """The serving stream: n_req f(R) n=1 requests, ngrid uniform over
    {64, 128, 192, 256} (cell 2 Mpc/h), |fR0| from {1e-4, 1e-5, 1e-6},
    seeded lognormal densities with sigma_g ~ U[0.8, 1.6] (request i: seed
    2000 + i), each solved to relative residual 1e-8."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([64, 128, 192, 256], n_req)
    fr0s = -rng.choice([1e-4, 1e-5, 1e-6], n_req)
    sigmas = rng.uniform(0.8, 1.6, n_req)
    cfgs = [M.Config(N=int(N), model=M.MODEL_FR, fr_n=1, box=2.0 * int(N), omega_m=0.3, omega_l=0.7, aexpn=1.0,
                     fr0=float(f), precision=M.PREC_MIXED) for N, f in zip(sizes, fr0s)]
    t0 = time.perf_counter()
    dens = lognormal_set(cfgs, sigmas, 2000)
    log(f"[serve] {n_req} densities generated in {time.perf_counter() - t0:.1f} s")
    out = serve_run("serve", cfgs, [{}] * n_req, dens, max_active, rng)
    out["sizes"] = [64, 128, 192, 256]
    return out


def serving_mix(n_req, seed, max_active):
    """nDGP requests mixed with f(R) ones in the serving loop: request i is
    nDGP (normal branch, H0rc from {0.5, 1, 2}) for even i, f(R) n=1 (|fR0|
    from {1e-4, 1e-5, 1e-6}) for odd i; ngrid uniform over the power-of-two
    sizes {64, 128, 256} (cell 2 Mpc/h); sigma_g ~ U[0.8, 1.6] (seed 6000 +
    i).  f(R) requests stop at relative residual 1e-8, nDGP requests by the
    reference's stagnation exits (their residual levels off)."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([64, 128, 256], n_req)
    fr0s = -rng.choice([1e-4, 1e-5, 1e-6], n_req)
    h0rcs = rng.choice([0.5, 1.0, 2.0], n_req)
    sigmas = rng.uniform(0.8, 1.6, n_req)
    cfgs, kws = [], []
    for i in range(n_req):
        N = int(sizes[i])
        if i % 2 == 0:
            cfgs.append(M.Config(N=N, model=M.MODEL_DGP, branch=1, box=2.0 * N, omega_m=0.3, omega_l=0.7,
                                 aexpn=1.0, h0rc=float(h0rcs[i]), precision=M.PREC_MIXED))
            kws.append(dict(exit_rule=M.EXIT_REFERENCE, res_conv=1e-30, max_outer=40))
        else:
            cfgs.append(M.Config(N=N, model=M.MODEL_FR, fr_n=1, box=2.0 * N, omega_m=0.3, omega_l=0.7, aexpn=1.0,
                                 fr0=float(fr0s[i]), precision=M.PREC_MIXED))
            kws.append({})
    t0 = time.perf_counter()
    dens = lognormal_set(cfgs, sigmas, 6000)
    log(f"[serve-mix] {n_req} densities generated in {time.perf_counter() - t0:.1f} s")
    out = serve_run("serve-mix", cfgs, kws, dens, max_active, rng)
    out["sizes"] = [64, 128, 256]
    out["models"] = "nDGP (even i) + f(R) (odd i)"
    return out


def profiled_traffic():
    """HBM bytes per launch of the roofline kernel, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled by a copy probe there)."""
    path = os.path.join(ROOT, "profiles", "fr512_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def copy_probe(gib):
    """A device copy of known size, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE."""
    import torch
    n = int(gib * (1 << 30) // 8)
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fr512", choices=sorted(WORKLOADS))
    ap.add_argument("--inner", type=int, default=0,
                    help="inner FP32 V-cycles per outer FP64 iteration (0: chosen per outer iteration)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-cycles", type=int, default=2)
    ap.add_argument("--copy-probe", type=float, default=0.0, help="only run a device copy of this many GiB")
    ap.add_argument("--serve-requests", type=int, default=512, help="requests of the serving stream (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-mix-requests", type=int, default=96,
                    help="requests of the nDGP + f(R) serving mix (0: skip)")
    ap.add_argument("--dgp-n", type=int, default=512, help="grid of the nDGP comparison (0: skip)")
    ap.add_argument("--pm-n", type=int, default=512, help="grid (and particles per side) of the PM step (0: skip)")
    args = ap.parse_args()
    # the CPU baseline's OpenMP runtimes read these when they load
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each solve runs on one device); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    # The JSON line goes to the original stdout; everything else written to
    # fd 1 (the Fortran runtime of the CPU baseline prints there and flushes
    # at exit) is sent to stderr so that the JSON line stays last.
    json_out = os.fdopen(os.dup(1), "w")
    sys.stdout.flush()
    os.dup2(2, 1)

    N, box, fr0, sigma, seed = WORKLOADS[args.workload]
    log(f"[bench] workload {args.workload}: N={N} box={box} fR0={fr0} sigma_g={sigma} seed={seed}")
    Smix = make_solver(N, box, fr0, M.PREC_MIXED)
    Smix.make_lognormal_density(seed, sigma)
    log("[bench] density generated")
    # the timed solves run without the library's per-kernel event profiler
    # (two events around every launch add 2-3 % to a mixed solve); the
    # per-kernel classes and the roofline come from PROF_STEPS profiled
    # solves of the same workload right after them
    reps, walls, total, _ = timed_solves(Smix, args.steps, args.warmup, False, args.inner)
    _, _, _, prof = timed_solves(Smix, PROF_STEPS, 0, True, args.inner)
    mix_ms = [r["seconds"] * 1e3 for r in reps]
    log(f"[bench] mixed solves: {['%.2f' % m for m in mix_ms]} ms, outer={reps[-1]['outer']} "
        f"inner={reps[-1]['cycles']} relres={reps[-1]['res_final'] / reps[-1]['res0']:.2e}")
    u_mix = Smix.get_field()
    kpk, pk_mix, _ = Smix.power_spectrum(2)

    fp64 = None
    if not args.no_fp64:
        S64 = make_solver(N, box, fr0, M.PREC_FP64)
        delta = Smix.get_density(np.float64)
        S64.set_density(delta)
        reps64, walls64, total64, _ = timed_solves(S64, args.steps, min(args.warmup, 1), False, 1)
        _, _, _, prof64 = timed_solves(S64, PROF_STEPS, 0, True, 1)
        u64 = S64.get_field()
        _, pk64, _ = S64.power_spectrum(2)
        ms64 = [r["seconds"] * 1e3 for r in reps64]
        log(f"[bench] fp64 solves: {['%.2f' % m for m in ms64]} ms, cycles={reps64[-1]['cycles']}")
        fu = np.abs(u_mix - u64)
        ratio = pk_mix / np.where(pk64 > 0, pk64, 1.0) - 1.0
        fp64 = {
            "ms_median": float(np.median(ms64)),
            "cycles": reps64[-1]["cycles"],
            "rel_residual": reps64[-1]["res_final"] / reps64[-1]["res0"],
            "mixed_vs_fp64": {
                "u_rel_linf": float(fu.max() / np.abs(u64).max()),
                "u_rel_l2": float(np.sqrt(np.mean(fu ** 2)) / np.sqrt(np.mean(u64 ** 2))),
                "fifth_force_pk_max_abs_ratio_minus_1": float(np.max(np.abs(ratio[pk64 > 0]))),
            },
            "smoother_avg_ms": prof64["smooth"][1] / max(prof64["smooth"][0], 1),
            "smoother_gbs": prof64["smooth"][2] / max(prof64["smooth"][1], 1e-9) / 1e6,
        }
        delta64 = delta
        del S64
    else:
        delta64 = Smix.get_density(np.float64)

    dgp = None
    if args.dgp_n > 0:
        dgp = dgp_compare(args.dgp_n, float(args.dgp_n), 1.2, 1005, 1)

    pm = None
    if args.pm_n > 0:
        pm = pm_step(args.pm_n, 77, 2)

    serving = None
    if args.serve_requests > 0:
        serving = serving_stream(args.serve_requests, 4000, args.serve_slots)
    serving_mixed = None
    if args.serve_mix_requests > 0:
        serving_mixed = serving_mix(args.serve_mix_requests, 4100, args.serve_slots)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_fp64_solve_baseline(delta64, N, box, fr0)
        ref = cpu_reference_baseline(delta64.astype(np.float32), N, box, fr0,
                                     fp64["cycles"] if fp64 else reps[-1]["outer"], args.cpu_cycles)
        if cpu is None:
            cpu = ref
        elif ref is not None:
            cpu["reference_fp32_cycles"] = ref

    sm = prof["smooth"]
    avg_ms = sm[1] / max(sm[0], 1)
    bytes_per_launch = sm[2] / max(sm[0], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    ms_step = float(np.median(mix_ms))
    lat = np.array(walls) * 1e3
    out = {
        "metric": "f(R) solve time to rel. residual 1e-8 @512^3 (mixed vs FP64); p50/p99 latency",
        "value": ms_step / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp32 smoothing/transfer + fp64 residual/update (mixed)",
        "data": "synthetic (seeded lognormal density generated on the GPU)",
        "config": {
            "workload": f"f(R) Hu-Sawicki n=1 |fR0|={abs(fr0):g} Om=0.3 a=1, {N}^3, box {box:g} Mpc/h, "
                        f"lognormal sigma_g={sigma} seed={seed}, rel. residual 1e-8 from u=u_bg",
            "N": N,
            "inner_cycles_per_outer": args.inner if args.inner > 0 else "auto",
        },
        "latency_ms": ({"p50": serving["delivered_p50_ms"], "p99": serving["delivered_p99_ms"],
                        "source": "serving stream (Poisson arrivals), submit -> field in the caller's array"}
                       if serving else None),
        "solve_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "serving": serving,
        "serving_dgp_fr_mix": serving_mixed,
        "dgp": dgp,
        "pm_step": pm,
        "mixed": {
            "outer_iterations": reps[-1]["outer"],
            "inner_vcycles": reps[-1]["cycles"],
            "rel_residual": reps[-1]["res_final"] / reps[-1]["res0"],
            "kernel_ms": {k: v[1] / PROF_STEPS for k, v in prof.items()},
        },
        "fp64": fp64,
        "speedup_mixed_vs_fp64": (fp64["ms_median"] / ms_step) if fp64 else None,
        "roofline": {
            "kernel": "level-0 smoother colour pass (mixed, FP32 delta form: k_colour_pass<DfSmoothOp<1, true>>)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic() if args.workload == "fr512" else None,
            "traffic_source": "profiles/fr512_hbm_traffic.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-scaled)",
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_ms,
            "launches": sm[0],
            "measured": f"HIP events around each launch of the kernel, on its stream, in {PROF_STEPS} profiled solves "
                        "of the same workload run right after the timed ones (which run without the events)",
        },
        "cpu_baseline": cpu,
    }
    json_out.write(json.dumps(out) + "\n")
    json_out.flush()


if __name__ == "__main__":
    main()
