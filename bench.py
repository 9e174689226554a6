"""Polarization-solve benchmark on one MI355X.

Workload: 15,000 polarizable AMOEBA sites uniform in a 33 A sphere (seeded
chain topology, polarization groups of 3-12 sites, p-scale 0/0/0.4/0.8), D and
P dipole sets.  One step = one complete polarization solve as after a geometry
update: static multipole field (FP64), both dipole sets from a cold start to
the tolerance, polarization energy.  Inputs stay resident in HBM; the field is
passed as a device pointer and the dipoles stay on the device.

Reported: solves/s of the mixed solver (FP32 T matvec inside FP64 iterative
refinement), the same for the plain FP64 CG, p50/p99 latency, the energy
error of both against the FP64 oracle's golden value (tight solve), live
timing of the dominant kernel (the FP32 on-the-fly pair kernel) against the
FP32 peak, and the oracle's CPU rate on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s @15k sites, mixed vs FP64 at ≤1e-9 rel E_pol err; p50/p99 solve latency"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
N_SITES, RADIUS, SEED = 15000, 33.0, 2


def pair_flops(n, m):
    """Ordered-pair on-the-fly matvec: N(N-1)(29 + 18m) FLOP per launch."""
    return n * (n - 1) * (29 + 18 * m)


def incore_bytes(n, m, s):
    """Packed symmetric T streamed once per matvec plus the vectors in and out:
    [N(N-1)/2 * 6 + N] * s + 2 * 3N * m * s bytes."""
    return (n * (n - 1) // 2 * 6 + n) * s + 2 * 3 * n * m * s


def profile_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "pair_kernels_sphere15000.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch") and kernel in line:
            return float(line.split()[-1])
    return None


def run_incore(spec, repeat):
    """Packed in-core FP32 T (two columns) streamed from HBM: mean device time per launch."""
    from openmmpol_amd import synth
    s = synth.make_system(spec)
    x = np.random.default_rng(1).normal(size=(2, s.pol_atoms, 3))
    s.profile_kernels(True)
    s.kernel_stats(reset=True)
    _, ms = s.tmatvec(x, kind="incore32", repeat=repeat)
    calls, kms = s.kernel_stats(reset=True)["incore32"]
    s.terminate()
    return ms, kms / max(calls, 1)


def load_golden():
    path = os.path.join(ROOT, "tests", "golden", f"sphere{N_SITES}_seed{SEED}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def run_mode(spec, mode, tol, steps, warmup, d_zero_ptr):
    from openmmpol_amd import synth
    s = synth.make_system(spec)
    s.set_precision(mode)
    s.set_solver_tolerance(tol, 500)
    # first solve builds device state (upload, scratch) outside the timed region
    s.reset_guess(True)
    s.set_external_field_device(d_zero_ptr)
    s.get_polelec_energy()
    for _ in range(warmup):
        s.reset_guess(True)
        s.set_external_field_device(d_zero_ptr)
        s.get_polelec_energy()
    s.profile_kernels(True)
    s.kernel_stats(reset=True)
    lat = []
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        s.reset_guess(True)                   # drop guess and static field (new geometry)
        s.set_external_field_device(d_zero_ptr)
        e = s.get_polelec_energy()            # returns after the device work completed
        lat.append((time.perf_counter() - t) * 1e3)
        iters.append(s.last_iterations)
    elapsed = time.perf_counter() - t0
    status = s.last_status
    stats = s.kernel_stats(reset=True)
    s.profile_kernels(False)
    ipd = s.ipd.copy()
    s.terminate()
    return dict(elapsed=elapsed, lat=np.array(lat), epol=e, ipd=ipd, iters=iters[-1], stats=stats,
                status=status)


def cpu_baseline(spec, fp64_iters):
    """Oracle (FP64 C/OpenMP restatement of the reference) on the host cores:
    static field once + 8 T matvecs of the same system, extrapolated to one
    full D+P CG solve with the FP64 iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as oa
    from openmmpol_amd import synth
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    o = oa.OracleSystem(synth.oracle_prep(spec))
    t = time.perf_counter()
    o.m2d()
    t_rhs = time.perf_counter() - t
    x = np.random.default_rng(0).normal(size=(o.npol, 3))
    nmv = 8
    t = time.perf_counter()
    for _ in range(nmv):
        x = 1e-3 * o.tmatvec(x)
    t_mv = (time.perf_counter() - t) / nmv
    n_mv = int(fp64_iters[0]) + int(fp64_iters[1]) + 2
    t_solve = t_rhs + n_mv * t_mv
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": int(oa.lib().oracle_num_threads()),
            "kind": "port",
            "sample": (f"oracle FP64 on {cores} host cores: static field once ({t_rhs:.2f} s) + {nmv} T matvecs "
                       f"({t_mv * 1e3:.0f} ms each) of the 15k-site system, extrapolated to one D+P CG solve of "
                       f"{n_mv} matvecs")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tol", type=float, default=3e-10,
                    help="solver tolerance on the preconditioned residual rms")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed-only", action="store_true")
    ap.add_argument("--no-incore", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    from openmmpol_amd import synth
    hip = ctypes.CDLL("libamdhip64.so")
    spec = synth.sphere_system(N_SITES, RADIUS, seed=SEED)
    n = len(spec["pol"])
    d_zero = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_zero), ctypes.c_size_t(3 * n * 8)) == 0
    assert hip.hipMemset(d_zero, 0, ctypes.c_size_t(3 * n * 8)) == 0
    print(f"[bench] system ready: {n} sites; running mixed", file=sys.stderr, flush=True)

    mixed = run_mode(spec, "mixed", args.tol, args.steps, args.warmup, d_zero.value)
    print(f"[bench] mixed: {args.steps / mixed['elapsed']:.2f} solves/s", file=sys.stderr, flush=True)
    fp64 = None
    if not args.mixed_only:
        fp64 = run_mode(spec, "fp64", args.tol, max(3, args.steps // 4), 1, d_zero.value)
        print(f"[bench] fp64: {len(fp64['lat']) / fp64['elapsed']:.2f} solves/s", file=sys.stderr, flush=True)
    hip.hipFree(d_zero)
    incore = None
    if not args.no_incore:
        incore = run_incore(spec, 20)
        print(f"[bench] incore32 matvec: {incore[1]:.3f} ms per stream launch", file=sys.stderr, flush=True)

    golden = load_golden()
    ref_e = golden["cases"]["zero"]["epol"] if golden else None

    gpath = os.path.join(ROOT, "tests", "golden", f"sphere{N_SITES}_seed{SEED}_ipd_zero.npy")
    gipd = np.load(gpath) if os.path.exists(gpath) else None

    def mu_err(ipd):
        return float(np.abs(ipd - gipd).max()) if gipd is not None else None
    value = args.steps / mixed["elapsed"]
    calls, ms = mixed["stats"]["otf32"]
    avg_ms = ms / max(calls, 1)
    flops = pair_flops(n, 2)
    achieved = flops / (avg_ms * 1e-3) / 1e12 if calls else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "15000-site AMOEBA sphere (R 33 A, seed 2), D+P dipoles, static field + "
                               "cold-start solve + energy per step, mixed FP32 matvec / FP64 refinement",
                   "n_sites": n, "tolerance": args.tol, "external_field": "zero", "parallelism": "single GPU"},
        "p50_ms": float(np.percentile(mixed["lat"], 50)),
        "p99_ms": float(np.percentile(mixed["lat"], 99)),
        "iterations_mixed": {"outer_fp64": mixed["iters"][2], "inner_fp32_total": mixed["iters"][3]},
        "epol_mixed": mixed["epol"],
        "epol_oracle_golden": ref_e,
        "epol_rel_err_mixed": abs(mixed["epol"] - ref_e) / abs(ref_e) if ref_e else None,
        "mu_maxabs_err_mixed": mu_err(mixed["ipd"]),
        "roofline": {"bound": "valu", "kernel": "k_tmv_pairs<float,2,AMOEBA> (on-the-fly FP32 pair kernel)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS if achieved else None,
                     "traffic": profile_traffic("k_tmv_pairs<float,2,AMOEBA>"),
                     "launches": calls, "avg_ms": avg_ms, "flop_per_launch": flops},
        "kernel_time_ms": {k: v[1] for k, v in mixed["stats"].items()},
        "kernel_launches": {k: v[0] for k, v in mixed["stats"].items()},
    }
    if incore is not None:
        b = incore_bytes(n, 2, 4)
        out["incore_fp32_matvec"] = {
            "kernel": "k_tpack_mv<float,2> (packed symmetric T, 64x64 site tiles)", "bound": "hbm",
            "avg_ms_stream_kernel": incore[1], "avg_ms_with_finalize": incore[0], "bytes_per_launch": b,
            "achieved": b / (incore[1] * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": b / (incore[1] * 1e-3) / 1e9 / PEAK_HBM_GBS}
    if fp64 is not None:
        v64 = len(fp64["lat"]) / fp64["elapsed"]
        c64, ms64 = fp64["stats"]["otf64"]
        out.update({
            "fp64_value": v64,
            "speedup_vs_fp64": value / v64,
            "fp64_p50_ms": float(np.percentile(fp64["lat"], 50)),
            "fp64_p99_ms": float(np.percentile(fp64["lat"], 99)),
            "iterations_fp64": {"D": fp64["iters"][0], "P": fp64["iters"][1]},
            "fp64_stop": "converged" if fp64["status"] == 0 else "reference small-step exit (|p.Tp| < eps_rp)",
            "epol_fp64": fp64["epol"],
            "epol_rel_err_fp64": abs(fp64["epol"] - ref_e) / abs(ref_e) if ref_e else None,
            "mu_maxabs_err_fp64": mu_err(fp64["ipd"]),
            "mu_maxabs_mixed_vs_fp64": float(np.abs(mixed["ipd"] - fp64["ipd"]).max()),
            "fp64_pair_kernel": {"launches": c64, "avg_ms": ms64 / max(c64, 1),
                                 "achieved_tflops": pair_flops(n, 1) / (ms64 / max(c64, 1) * 1e-3) / 1e12
                                 if c64 else None, "peak": PEAK_FP64_TFLOPS},
        })
        if not args.no_cpu_baseline:
            print("[bench] cpu baseline", file=sys.stderr, flush=True)
            out["cpu_baseline"] = cpu_baseline(spec, fp64["iters"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
