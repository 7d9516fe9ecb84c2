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
    ap.add_argument("--inner", type=int, default=1, help="inner FP32 V-cycles per outer FP64 iteration")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-cycles", type=int, default=2)
    ap.add_argument("--copy-probe", type=float, default=0.0, help="only run a device copy of this many GiB")
    args = ap.parse_args()
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
    reps, walls, total, prof = timed_solves(Smix, args.steps, args.warmup, True, args.inner)
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
        reps64, walls64, total64, prof64 = timed_solves(S64, args.steps, min(args.warmup, 1), True, 1)
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
        delta32 = delta.astype(np.float32)
        del S64
    else:
        delta32 = Smix.get_density(np.float32)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_reference_baseline(delta32, N, box, fr0, fp64["cycles"] if fp64 else reps[-1]["outer"],
                                     args.cpu_cycles)

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
            "inner_cycles_per_outer": args.inner,
        },
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "mixed": {
            "outer_iterations": reps[-1]["outer"],
            "inner_vcycles": reps[-1]["cycles"],
            "rel_residual": reps[-1]["res_final"] / reps[-1]["res0"],
            "kernel_ms": {k: v[1] / max(args.steps, 1) for k, v in prof.items()},
        },
        "fp64": fp64,
        "speedup_mixed_vs_fp64": (fp64["ms_median"] / ms_step) if fp64 else None,
        "roofline": {
            "kernel": "level-0 smoother colour pass (mixed, FP32 delta form: k_colour_pass<DfSmoothOp<1>>)",
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
        },
        "cpu_baseline": cpu,
    }
    json_out.write(json.dumps(out) + "\n")
    json_out.flush()


if __name__ == "__main__":
    main()
