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
PEAK_BF16_TFLOPS = 2500.0  # dense matrix cores
PEAK_HBM_GBS = 8000.0
N_SITES, RADIUS, SEED = 15000, 33.0, 2
PROF_EVERY = 4


def sym_pair_flops(n, m):
    """Pair-once on-the-fly matvec (each unordered pair evaluated once, both
    directions accumulated): N(N-1)/2 (29 + 36m) FLOP per launch -- 29 for the
    geometry and damping of the pair, 18 per column and direction."""
    return n * (n - 1) // 2 * (29 + 36 * m)


def grad_flops(npol, nmm):
    """FP64 FLOP per gradient (FMA = 2, exponentials of damped pairs not
    counted).  Below 5,000 polarizable sites: the static-multipole sweep 135 per
    (polarizable target, MM source) pair, the MM-target sweep 166 per (MM
    target, polarizable source) pair, the dipole-dipole sweep 113 per ordered
    polarizable pair.  From 5,000, where each (polarizable, MM) pair of the
    static part is evaluated once for both ends and each unordered polarizable
    pair once in the dipole-dipole part: 200 per (polarizable, MM) pair (the
    135 of the polarizable end, 58 for the field and field gradient of the
    dipoles at the MM end, 6 for the two ends' sums) and 113 per unordered
    polarizable pair."""
    if npol >= 5000:
        return 200 * npol * nmm + 113 * (npol * (npol - 1) // 2)
    return 135 * npol * nmm + 113 * npol * npol + 166 * nmm * npol


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
    """Packed in-core FP32 T (two columns) streamed from HBM: mean device time per
    launch of the whole matvec and of the stream kernel alone."""
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


def run_mode(spec, mode, tol, steps, warmup, d_zero_ptr, host_steps=0):
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
    # HIP events bracket every PROF_EVERY-th pair-kernel launch of the timed
    # region: the two event records cost ~12 us of idle GPU per bracketed launch
    s.profile_kernels(True, every=PROF_EVERY)
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
    # the same step through the drop-in host-pointer entry (ommp_set_external_field:
    # the field copied in from host memory, the dipoles back after the solve)
    host = None
    if host_steps:
        ext = np.zeros((s.pol_atoms, 3))
        s.reset_guess(True)
        s.set_external_field(ext)
        s.get_polelec_energy()
        t0 = time.perf_counter()
        for _ in range(host_steps):
            s.reset_guess(True)
            s.set_external_field(ext)
            s.get_polelec_energy()
            _ = s.ipd[0, 0, 0]
        host = (time.perf_counter() - t0) / host_steps
    s.terminate()
    return dict(elapsed=elapsed, lat=np.array(lat), epol=e, ipd=ipd, iters=iters[-1], stats=stats,
                status=status, host=host)


class DeviceBuffers:
    """Raw device allocations through the HIP runtime (inputs resident in HBM)."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(self.ct.c_void_p(p), a.ctypes.data_as(self.ct.c_void_p),
                                  self.ct.c_size_t(a.nbytes), 1) == 0
        return p

    def alloc(self, nbytes):
        p = self.ct.c_void_p()
        assert self.hip.hipMalloc(self.ct.byref(p), self.ct.c_size_t(max(nbytes, 8))) == 0
        self.ptrs.append(p)
        return p.value

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


MULTI_NRHS, MULTI_QSEED = 64, 21


def run_multi(tol, steps, bufs):
    """64 QM-region fields on the 15k-site system with a 6 A cavity: one batched
    call (all columns share every launch) against the same fields solved one
    after another with the single-field entry (each warm-started from the
    previous one), mixed and plain FP64."""
    from openmmpol_amd import synth
    spec = synth.sphere_system(N_SITES, RADIUS, seed=SEED, cavity=6.0)
    charges = synth.qm_charge_sets(MULTI_NRHS, MULTI_QSEED)
    s = synth.make_system(spec)
    cpol = spec["cmm"][np.asarray(s.polar_mm)]
    ext = np.stack([synth.point_charge_field(cpol, q[:, :3], q[:, 3]) for q in charges])
    n = s.pol_atoms
    d_ext = bufs.upload(ext)
    d_ipd = bufs.alloc(MULTI_NRHS * 2 * 3 * n * 8)
    s.set_precision("mixed")
    s.set_solver_tolerance(tol, 500)
    # the previous solve: the undisplaced QM field; its dipoles are every column's guess
    s.set_external_field_device(d_ext)
    s.get_polelec_energy()
    ep = s.set_external_fields_multi_device(MULTI_NRHS, d_ext, d_ipd)  # warm-up (allocates batch state)
    t = time.perf_counter()
    for _ in range(steps):
        ep = s.set_external_fields_multi_device(MULTI_NRHS, d_ext, d_ipd)
    t_multi = (time.perf_counter() - t) / steps
    it_multi = s.last_iterations
    ipd_all = np.zeros((MULTI_NRHS, 2, n, 3))
    bufs.hip.hipMemcpy(ipd_all.ctypes.data_as(bufs.ct.c_void_p), bufs.ct.c_void_p(d_ipd),
                       bufs.ct.c_size_t(ipd_all.nbytes), 2)

    def sequential(mode, count):
        s.set_precision(mode)
        s.reset_guess(False)
        s.set_external_field_device(d_ext)  # same starting point as the batched call
        s.get_polelec_energy()
        e = []
        t0 = time.perf_counter()
        for k in range(count):
            s.set_external_field_device(d_ext + k * 3 * n * 8)
            e.append(s.get_polelec_energy())
        return (time.perf_counter() - t0) / count, np.array(e)

    # the block matvec of the inner solves alone, 128 columns: algorithmic FLOPs
    # N(N-1) x 18 x 128 (the 3x3 tensor times each column, every ordered pair);
    # the matrix cores run six bf16 products per pair of split operands over
    # tiles padded to 32-site multiples
    xb = np.random.default_rng(7).normal(0.0, 0.05, size=(128, n, 3))
    _, bmv_ms = s.block_tmatvec(xb, repeat=5)
    npad = -(-n // 32) * 32
    bmv = {"ms_per_launch": bmv_ms, "columns": 128,
           "algorithmic_tflops": n * (n - 1) * 18 * 128 / (bmv_ms * 1e-3) / 1e12,
           "mfma_bf16_tflops": 6 * npad * npad * 9 * 2 * 128 / (bmv_ms * 1e-3) / 1e12,
           "peak_bf16_tflops": PEAK_BF16_TFLOPS}
    bmv["mfma_frac"] = bmv["mfma_bf16_tflops"] / PEAK_BF16_TFLOPS
    t_seq_mixed, e_seq_mixed = sequential("mixed", MULTI_NRHS)
    t_seq_fp64, e_seq_fp64 = sequential("fp64", 16)
    s.terminate()
    gpath = os.path.join(ROOT, "tests", "golden", "multi15000_cavity.json")
    err_e, err_mu = None, None
    if os.path.exists(gpath):
        with open(gpath) as f:
            g = json.load(f)
        err_e = max(abs(ep[r["field"]] - r["epol"]) / abs(r["epol"]) for r in g["rhs"])
        # dipoles at the golden's sample sites of every golden field (field 0 is
        # its own guess; the others are solved by the batched kernels)
        err_mu = max(float(np.abs(ipd_all[r["field"]][:, g["sample"], :] - np.array(r["ipd_sample"])).max())
                     for r in g["rhs"])
    return {
        "workload": f"15000-site AMOEBA sphere with a 6 A cavity, {MULTI_NRHS} fields of a 30-charge QM region "
                    "displaced N(0, 0.1 A) per field; guess = dipoles of the undisplaced-field solve",
        "rhs_solves_per_s": MULTI_NRHS / t_multi, "ms_per_call": t_multi * 1e3,
        "inner_matvec": "FP32 tensors and p split in three bf16 parts each, the six products A_i X_j "
                        "(i + j <= 2) on the matrix cores (v_mfma_f32_16x16x32_bf16), FP32 accumulation "
                        "(FP32-class error); FP64 refinement residuals",
        "iterations": {"outer_fp64_max": it_multi[2], "inner_max": it_multi[3]},
        "block_matvec": bmv,
        "sequential_mixed_solves_per_s": 1.0 / t_seq_mixed,
        "sequential_fp64_solves_per_s": 1.0 / t_seq_fp64,
        "speedup_vs_sequential_mixed": t_seq_mixed / t_multi * MULTI_NRHS,
        "speedup_vs_sequential_fp64": t_seq_fp64 / t_multi * MULTI_NRHS,
        "epol_rel_err_vs_golden_max": err_e, "mu_maxabs_err_vs_golden_samples": err_mu,
        "epol_max_rel_diff_vs_sequential_fp64": float(np.max(np.abs(ep[:16] - e_seq_fp64) / np.abs(e_seq_fp64))),
    }


def run_serving(tol, nreq, seed, mean_gap_ms, max_active, bufs):
    """Poisson arrivals of SCF-like requests on systems of log-uniform size
    500-40,000 sites; the continuous-batching server against the same requests
    run one after another through the single-field entry in plain FP64."""
    from openmmpol_amd import Server, synth
    sizes, nsteps, arrivals, charges, make_spec = synth.serving_requests(nreq, seed, mean_gap_ms)
    specs = [make_spec(k) for k in range(nreq)]
    total = int(nsteps.sum())
    # ---- server
    systems = [synth.make_system(sp) for sp in specs]
    for x in systems:
        x.set_solver_tolerance(tol, 500)
    srv = Server(hbm_budget=200e9, max_active=max_active)
    t0 = time.perf_counter()
    ids = []
    for k in range(nreq):
        dt = t0 + arrivals[k] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(systems[k], charges[k]))
    srv.wait()
    makespan = time.perf_counter() - t0
    lat, e_srv, its = [], [], []
    for rid in ids:
        st, e, la, it = srv.result(rid)
        assert st == 0
        lat.extend(la.tolist())
        e_srv.append(e)
        its.append(it)
    its = np.concatenate(its)
    stats = srv.stats()
    srv.close()
    for x in systems:
        x.terminate()
    # ---- sequential plain FP64, one solve after another (fields precomputed in HBM)
    systems = [synth.make_system(sp) for sp in specs]
    fields = []
    for k in range(nreq):
        cpol = specs[k]["cmm"][np.asarray(systems[k].polar_mm)]
        f = np.stack([synth.point_charge_field(cpol, q[:, :3], q[:, 3]) for q in charges[k]])
        fields.append((bufs.upload(f), f.shape[1]))
    seq_lat, e_seq, seq_it = [], [], []
    t = time.perf_counter()
    for k in range(nreq):
        x = systems[k]
        x.set_solver_tolerance(tol, 500)
        ee = []
        d_f, npol = fields[k]
        for j in range(nsteps[k]):
            t1 = time.perf_counter()
            x.set_external_field_device(d_f + j * 3 * npol * 8)
            ee.append(x.get_polelec_energy())
            seq_lat.append((time.perf_counter() - t1) * 1e3)
            seq_it.append(x.last_iterations[:2])
        e_seq.append(np.array(ee))
    t_seq = time.perf_counter() - t
    for x in systems:
        x.terminate()
    # accuracy: the first and last solve of four requests (the two largest
    # among them) against the oracle's tight solves of the same fields
    # (tests/golden/make_serving_golden.py), and every solve of the server
    # against the same solve of the sequential plain FP64 CG
    gpath = os.path.join(ROOT, "tests", "golden", f"serving{nreq}_seed{seed}.json")
    err_srv = err_seq = None
    golden_cases = 0
    if os.path.exists(gpath):
        with open(gpath) as f:
            g = json.load(f)
        if g["nreq"] == nreq and g["seed"] == seed:
            err_srv = max(abs(e_srv[c["request"]][c["step"]] - c["epol"]) / abs(c["epol"]) for c in g["cases"])
            err_seq = max(abs(e_seq[c["request"]][c["step"]] - c["epol"]) / abs(c["epol"]) for c in g["cases"])
            golden_cases = len(g["cases"])
    diff_srv_seq = max(float(np.max(np.abs(a - b) / np.abs(b))) for a, b in zip(e_srv, e_seq))
    lat = np.array(lat)
    seq_lat = np.array(seq_lat)
    req_size = np.concatenate([np.full(nsteps[k], sizes[k]) for k in range(nreq)])
    # a request's first solve counts from its arrival (admission, upload and
    # static field included); later solves from the previous solve's completion
    first = np.concatenate([np.arange(nsteps[k]) == 0 for k in range(nreq)])
    by_size = {}
    for lo, hi in ((500, 2000), (2000, 10000), (10000, 40001)):
        sel = (req_size >= lo) & (req_size < hi)
        if sel.any():
            later = sel & ~first
            by_size[f"{lo}-{min(hi, 40000)}"] = {
                "solves": int(sel.sum()),
                "server_p50_ms": float(np.percentile(lat[sel], 50)),
                "server_p99_ms": float(np.percentile(lat[sel], 99)),
                "server_first_solve_max_ms": float(lat[sel & first].max()),
                "server_later_solves_p50_ms": float(np.percentile(lat[later], 50)),
                "server_later_solves_p99_ms": float(np.percentile(lat[later], 99)),
                "sequential_fp64_p50_ms": float(np.percentile(seq_lat[sel], 50)),
                "sequential_fp64_p99_ms": float(np.percentile(seq_lat[sel], 99))}
    return {
        "workload": (f"{nreq} requests, sizes log-uniform 500-40000 sites (seed {seed}), 10-20 solves each "
                     f"({total} solves), 30-charge QM field perturbed N(0, 0.02 A) per solve, Poisson arrivals "
                     f"every {mean_gap_ms} ms on average, up to {max_active} requests admitted"),
        "solves_per_s": total / makespan, "makespan_s": makespan,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "sequential_fp64_solves_per_s": total / t_seq,
        "sequential_fp64_p50_ms": float(np.percentile(seq_lat, 50)),
        "sequential_fp64_p99_ms": float(np.percentile(seq_lat, 99)),
        "latency_by_request_size": by_size,
        "speedup_vs_sequential_fp64": t_seq / makespan,
        "epol_max_rel_err_server_vs_oracle": err_srv,
        "epol_max_rel_err_sequential_fp64_vs_oracle": err_seq,
        "oracle_golden_solves": golden_cases,
        "epol_max_rel_diff_server_vs_sequential_fp64": diff_srv_seq,
        "iterations_server_mean": {"outer_fp64": float(its[:, 0].mean()), "inner_fp32": float(its[:, 1].mean())},
        "iterations_sequential_fp64_mean": float(np.mean(seq_it)),
        "server_stats": stats,
    }


def run_large(tol, bufs):
    """60,000 sites in a 52 A sphere: mixed solves with the on-the-fly FP32
    matvec and with the packed FP32 T in HBM (43 GB), plain FP64 CG, and the
    packed-T stream rate."""
    from openmmpol_amd import synth
    n_sites, radius, seed = 60000, 52.0, 3
    spec = synth.sphere_system(n_sites, radius, seed=seed)
    s = synth.make_system(spec)
    n = s.pol_atoms
    d_zero = bufs.alloc(3 * n * 8)
    bufs.hip.hipMemset(bufs.ct.c_void_p(d_zero), 0, bufs.ct.c_size_t(3 * n * 8))
    s.set_solver_tolerance(tol, 500)

    def solve(mode, matv, steps):
        s.set_precision(mode)
        s.set_default_matv(matv)
        out = []
        for k in range(steps + 1):  # the first (untimed) builds device state / packed T
            t = time.perf_counter()
            s.reset_guess(True)
            s.set_external_field_device(d_zero)
            e = s.get_polelec_energy()
            if k:
                out.append(time.perf_counter() - t)
        return float(np.mean(out)), e, s.ipd.copy(), s.last_iterations

    t_otf, e_otf, ipd_otf, it_otf = solve("mixed", "direct", 2)
    t_inc, e_inc, ipd_inc, it_inc = solve("mixed", "incore", 2)
    x = np.random.default_rng(1).normal(size=(2, n, 3))
    s.profile_kernels(True)
    s.kernel_stats(reset=True)
    _, ms_inc = s.tmatvec(x, kind="incore32", repeat=5)
    st = s.kernel_stats(reset=True)
    s.profile_kernels(False)
    t_64, e_64, ipd_64, it_64 = solve("fp64", "direct", 1)
    s.terminate()
    b = incore_bytes(n, 2, 4)
    stream_ms = st["incore32"][1] / max(st["incore32"][0], 1)
    res = {
        "workload": f"{n_sites}-site AMOEBA sphere (R {radius} A, seed {seed}), static field + cold-start D+P "
                    "solve + energy",
        "mixed_otf_solves_per_s": 1.0 / t_otf, "mixed_incore_solves_per_s": 1.0 / t_inc,
        "fp64_solves_per_s": 1.0 / t_64, "speedup_mixed_otf_vs_fp64": t_64 / t_otf,
        "iterations": {"mixed_otf": it_otf, "mixed_incore": it_inc, "fp64": it_64[:2]},
        "incore_fp32_stream": {"packed_T_bytes": b, "avg_ms": stream_ms,
                               "achieved_GBs": b / (stream_ms * 1e-3) / 1e9, "peak_GBs": PEAK_HBM_GBS},
        "mu_maxabs_mixed_otf_vs_incore": float(np.abs(ipd_otf - ipd_inc).max()),
    }
    gpath = os.path.join(ROOT, "tests", "golden", f"sphere{n_sites}_seed{seed}.json")
    if os.path.exists(gpath):
        with open(gpath) as f:
            g = json.load(f)
        smp = g["sample"]
        for label, e, ipd in (("mixed_otf", e_otf, ipd_otf), ("mixed_incore", e_inc, ipd_inc), ("fp64", e_64, ipd_64)):
            res[f"epol_rel_err_{label}"] = abs(e - g["epol"]) / abs(g["epol"])
            res[f"mu_sample_maxabs_err_{label}"] = float(np.abs(ipd[:, smp, :] - np.array(g["ipd_sample"])).max())
    return res


def run_grad(spec, tol, repeat, cpu):
    """d E_pol / d r of every site (ommp_polelec_geomgrad) after a converged
    FP64 solve: wall time per call (copy-back of the 3N gradient included) and
    live event timing of the three FP64 pair sweeps.  With cpu, the oracle's
    gradient of the same dipoles is timed once on the host cores (the full
    15,000 sites, about 10 s of CPU work) and compared with the GPU's."""
    from openmmpol_amd import synth
    s = synth.make_system(spec)
    s.set_precision("fp64")
    s.set_solver_tolerance(tol, 500)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    ipd = s.ipd.copy()
    g = s.polelec_geomgrad()  # first call builds the gradient's lists and buffers
    s.profile_kernels(True)
    s.kernel_stats(reset=True)
    t = time.perf_counter()
    for _ in range(repeat):
        g = s.polelec_geomgrad()
    wall = (time.perf_counter() - t) / repeat
    calls, ms = s.kernel_stats(reset=True)["geomgrad"]
    s.profile_kernels(False)
    npol, nmm = s.pol_atoms, s.mm_atoms
    s.terminate()
    sweep_ms = ms / max(calls, 1)
    flops = grad_flops(npol, nmm)
    res = {"workload": f"{nmm}-site AMOEBA sphere, gradient of E_pol after a converged FP64 solve",
           "ms_per_gradient": wall * 1e3, "gradients_per_s": 1.0 / wall,
           "pair_sweeps_avg_ms": sweep_ms, "launches": calls, "flop_per_gradient": flops,
           "achieved_tflops": flops / (sweep_ms * 1e-3) / 1e12 if calls else None,
           "peak_tflops": PEAK_FP64_TFLOPS, "bound": "valu (FP64)",
           "sum_rule_rel": float(np.abs(g.sum(axis=0)).max() / np.abs(g).max())}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_api as oa
        o = oa.OracleSystem(synth.oracle_prep(spec))
        t = time.perf_counter()
        go = o.polelec_geomgrad(ipd)
        t_cpu = time.perf_counter() - t
        res["cpu_baseline"] = {"value": 1.0 / t_cpu, "unit": "gradients/s", "cores": oa.lib().oracle_num_threads(),
                               "kind": "port", "sample": f"one oracle gradient of the same {nmm} sites and dipoles"}
        res["max_abs_diff_vs_oracle_rel"] = float(np.abs(g - go).max() / np.abs(go).max())
    return res


def fmm_m2l_flops(cmm, p):
    """Complex multiply-adds of the M2L stage as 8 FLOP each: for every level
    l >= 2 and every occupied target cell, (p+1)^4 per occupied source cell in
    its interaction list (children of the parent's neighbours that are not
    neighbours).  Same octree as the library (leaves >= 7 A)."""
    lo, hi = cmm.min(0), cmm.max(0)
    side = float((hi - lo).max()) * (1 + 1e-9) + 1e-6
    origin = lo - 0.5 * (side - (hi - lo))
    leaf_min = 7.0 * 1.8897261245650
    L = 0
    while L < 7 and side / 2 ** (L + 1) >= leaf_min:
        L += 1
    total = 0
    for lev in range(2, L + 1):
        ns = 2 ** lev
        idx = np.clip(np.floor((cmm - origin) / (side / ns)).astype(int), 0, ns - 1)
        occ = np.zeros((ns, ns, ns), dtype=bool)
        occ[idx[:, 2], idx[:, 1], idx[:, 0]] = True
        cells = np.argwhere(occ)
        for c in cells:
            base = 2 * (c // 2) - 2
            lo3, hi3 = np.maximum(base, 0), np.minimum(base + 5, ns - 1)
            sub = occ[lo3[0]:hi3[0] + 1, lo3[1]:hi3[1] + 1, lo3[2]:hi3[2] + 1]
            near = occ[max(c[0] - 1, 0):c[0] + 2, max(c[1] - 1, 0):c[1] + 2, max(c[2] - 1, 0):c[2] + 2]
            total += int(sub.sum()) - int(near.sum())
    return 8 * total * (p + 1) ** 4, L


def run_fmm(tol):
    """250,000 AMOEBA sites in an 84 A sphere: the fast-multipole T matvec
    (order 8, the reference's default) against the direct FP64 on-the-fly one
    (time, largest-component error), a complete FP64 solve with fast
    multipoles against the mixed direct solve of the same system, and the
    gradient of E_pol through the tree against the direct gradient (each after
    its own solve, so their difference includes that of the dipoles)."""
    from openmmpol_amd import synth
    n_sites, radius, seed, p = 250000, 84.0, 5, 8
    spec = synth.sphere_system(n_sites, radius, seed=seed)
    s = synth.make_system(spec)
    n = s.pol_atoms
    x = np.random.default_rng(4).normal(size=(2, n, 3))
    y_dir, ms_dir = s.tmatvec(x, kind="otf64", repeat=2)
    _, ms_dir32 = s.tmatvec(x, kind="otf32", repeat=2)
    s.enable_fmm(lmax=p, lmax_pol=p)
    y_fmm, ms_fmm = s.tmatvec(x, kind="fmm", repeat=5)
    err = float(np.abs(y_fmm - y_dir).max() / np.abs(y_dir).max())
    s.set_solver_tolerance(tol, 500)
    t = time.perf_counter()
    s.set_external_field(np.zeros((n, 3)))
    e_fmm = s.get_polelec_energy()
    t_fmm = time.perf_counter() - t
    it_fmm = s.last_iterations[:2]
    # gradient through the tree (first call builds the gradient's lists)
    g_fmm = s.polelec_geomgrad()
    s.profile_kernels(True)
    s.kernel_stats(reset=True)
    t = time.perf_counter()
    for _ in range(3):
        g_fmm = s.polelec_geomgrad()
    t_gfmm = (time.perf_counter() - t) / 3
    gcalls, gms = s.kernel_stats(reset=True)["geomgrad"]
    s.profile_kernels(False)
    s.disable_fmm()
    s.set_precision("mixed")
    t = time.perf_counter()
    s.set_external_field(np.zeros((n, 3)))
    e_dir = s.get_polelec_energy()
    t_dir = time.perf_counter() - t
    t = time.perf_counter()
    g_dir = s.polelec_geomgrad()
    t_gdir = time.perf_counter() - t
    s.terminate()
    flops, levels = fmm_m2l_flops(np.asarray(spec["cmm"]), p)
    return {"workload": f"{n_sites}-site AMOEBA sphere (R {radius} A, seed {seed}), two dipole columns",
            "order": p, "levels": levels,
            "fmm_matvec_ms": ms_fmm, "direct_fp64_matvec_ms": ms_dir, "direct_fp32_matvec_ms": ms_dir32,
            "speedup_vs_direct_fp64": ms_dir / ms_fmm, "matvec_rel_err": err,
            "m2l_flop_per_matvec": 2 * flops,
            "solve_fmm_s": t_fmm, "solve_direct_mixed_s": t_dir, "iterations_fmm": list(it_fmm),
            "epol_fmm": e_fmm, "epol_direct": e_dir, "epol_rel_diff": abs(e_fmm - e_dir) / abs(e_dir),
            "grad_fmm_ms": t_gfmm * 1e3, "grad_fmm_pair_sums_ms": gms / max(gcalls, 1),
            "grad_direct_ms": t_gdir * 1e3, "grad_speedup_vs_direct": t_gdir / t_gfmm,
            "grad_rel_diff_vs_direct": float(np.abs(g_fmm - g_dir).max() / np.abs(g_dir).max())}


def cpu_baseline(spec, tol):
    """Oracle (FP64 C/OpenMP restatement of the reference) on the host cores:
    one complete timed solve of the headline step -- static field, D and P
    CG solves from a cold start to the same tolerance with the reference's
    stopping rule, energy."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as oa
    from openmmpol_amd import synth
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    o = oa.OracleSystem(synth.oracle_prep(spec))
    t = time.perf_counter()
    o.set_external_field(np.zeros((o.npol, 3)), tol=tol)
    e = o.get_polelec_energy()
    t_solve = time.perf_counter() - t
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": int(oa.lib().oracle_num_threads()),
            "kind": "port", "epol": e, "iterations": {"D": int(o.iters[0]), "P": int(o.iters[1])},
            "sample": (f"one complete oracle FP64 solve of the 15k-site step (static field, D and P CG from a cold "
                       f"start to tol {tol:g}, energy) on {cores} host cores, timed: {t_solve:.2f} s")}


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
    ap.add_argument("--no-multi", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=16)
    ap.add_argument("--serving-gap-ms", type=float, default=10.0)
    ap.add_argument("--serving-slots", type=int, default=16)
    ap.add_argument("--incore-sites", default="15000",
                    help="comma-separated sphere sizes of the packed-T stream leg (density as the 15k sphere); "
                         "the bench line reports the first")
    ap.add_argument("--legs", default="headline,fp64,multi,serving,incore,large,grad,fmm,cpu",
                    help="comma-separated subset to run (profiling one leg); the default runs all")
    args = ap.parse_args()
    legs = set(args.legs.split(","))
    for flag, leg in ((args.no_cpu_baseline, "cpu"), (args.mixed_only, "fp64"), (args.no_incore, "incore"),
                      (args.no_multi, "multi"), (args.no_serving, "serving")):
        if flag:
            legs.discard(leg)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(one process drives one GPU)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    from openmmpol_amd import synth
    hip = ctypes.CDLL("libamdhip64.so")
    spec = synth.sphere_system(N_SITES, RADIUS, seed=SEED)
    n = len(spec["pol"])
    d_zero = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_zero), ctypes.c_size_t(3 * n * 8)) == 0
    assert hip.hipMemset(d_zero, 0, ctypes.c_size_t(3 * n * 8)) == 0
    print(f"[bench] system ready: {n} sites", file=sys.stderr, flush=True)

    mixed = None
    if "headline" in legs:
        mixed = run_mode(spec, "mixed", args.tol, args.steps, args.warmup, d_zero.value, host_steps=args.steps)
        print(f"[bench] mixed: {args.steps / mixed['elapsed']:.2f} solves/s", file=sys.stderr, flush=True)
    fp64 = None
    if "fp64" in legs:
        fp64 = run_mode(spec, "fp64", args.tol, max(3, args.steps // 4), 1, d_zero.value)
        print(f"[bench] fp64: {len(fp64['lat']) / fp64['elapsed']:.2f} solves/s", file=sys.stderr, flush=True)
    hip.hipFree(d_zero)
    bufs = DeviceBuffers()
    multi = None
    if "multi" in legs:
        multi = run_multi(args.tol, 3, bufs)
        print(f"[bench] multi-RHS: {multi['rhs_solves_per_s']:.1f} rhs solves/s", file=sys.stderr, flush=True)
    bufs.free()
    serving = None
    if "serving" in legs:
        serving = run_serving(args.tol, args.serving_requests, 5, args.serving_gap_ms, args.serving_slots, bufs)
        bufs.free()
        print(f"[bench] serving: {serving['solves_per_s']:.1f} solves/s", file=sys.stderr, flush=True)
    large = None
    if "large" in legs:
        large = run_large(args.tol, bufs)
        bufs.free()
        print(f"[bench] 60k sites: mixed {large['mixed_otf_solves_per_s']:.2f} solves/s", file=sys.stderr,
              flush=True)
    grad = None
    if "grad" in legs:
        grad = run_grad(spec, args.tol, 10, "cpu" in legs)
        print(f"[bench] gradient: {grad['ms_per_gradient']:.2f} ms per call", file=sys.stderr, flush=True)
    fmm = None
    if "fmm" in legs:
        fmm = run_fmm(args.tol)
        print(f"[bench] FMM matvec: {fmm['fmm_matvec_ms']:.2f} ms vs direct {fmm['direct_fp64_matvec_ms']:.2f} ms; "
              f"FMM gradient {fmm['grad_fmm_ms']:.2f} ms vs direct {fmm['grad_direct_ms']:.2f} ms",
              file=sys.stderr, flush=True)
    incore = None
    incore_extra = []
    if "incore" in legs:
        for k, ns in enumerate(int(v) for v in args.incore_sites.split(",")):
            sp = spec if ns == N_SITES else synth.sphere_system(ns, RADIUS * (ns / N_SITES) ** (1.0 / 3.0),
                                                                seed=SEED)
            r = run_incore(sp, 20)
            print(f"[bench] incore32 matvec at {ns} sites: {r[1]:.3f} ms per stream launch", file=sys.stderr,
                  flush=True)
            if k == 0:
                incore = r
            else:
                b = incore_bytes(ns, 2, 4)
                incore_extra.append({"n_sites": ns, "avg_ms_stream_kernel": r[1], "avg_ms_with_finalize": r[0],
                                     "bytes_per_launch": b, "achieved_GBs": b / (r[1] * 1e-3) / 1e9})

    golden = load_golden()
    ref_e = golden["cases"]["zero"]["epol"] if golden else None

    gpath = os.path.join(ROOT, "tests", "golden", f"sphere{N_SITES}_seed{SEED}_ipd_zero.npy")
    gipd = np.load(gpath) if os.path.exists(gpath) else None

    def mu_err(ipd):
        return float(np.abs(ipd - gipd).max()) if gipd is not None else None
    if mixed is None:
        # a subset of the legs (profiling): no headline figure
        out = {"legs": sorted(legs)}
        for key, val in (("multi_rhs", multi), ("serving", serving), ("large_60k", large), ("geomgrad", grad),
                         ("fmm_250k", fmm)):
            if val is not None:
                out[key] = val
        if incore is not None:
            out["incore_fp32_stream"] = [{"avg_ms_stream_kernel": incore[1], "avg_ms_with_finalize": incore[0]}] + \
                incore_extra
        print(json.dumps(out))
        return
    value = args.steps / mixed["elapsed"]
    calls, ms = mixed["stats"]["otf32"]
    avg_ms = ms / max(calls, 1)
    # at 15,000 sites both pair kernels are the pair-once sweep k_tmv_sym
    flops = sym_pair_flops(n, 2)
    achieved = flops / (avg_ms * 1e-3) / 1e12 if calls else None
    # the rate non-packed vector FMAs reach on this device (the build emits no
    # packed FP32 instructions; the published FP32 peak counts them)
    from openmmpol_amd._lib import lib as ommp_lib
    vfma = {"fp32_tflops": ommp_lib().ommp_amd_vector_fma_tflops(32),
            "fp64_tflops": ommp_lib().ommp_amd_vector_fma_tflops(64),
            "method": "8 independent FMA chains per thread, 8192 x 256 threads, no packed FP32 instructions"}
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
        "host_entry_solves_per_s": 1.0 / mixed["host"] if mixed.get("host") else None,
        "host_entry_note": "the same step through the drop-in host-pointer entry (field copied in from "
                           "pageable host memory, dipoles copied back); not the value",
        "iterations_mixed": {"outer_fp64": mixed["iters"][2], "inner_fp32_total": mixed["iters"][3]},
        "epol_mixed": mixed["epol"],
        "epol_oracle_golden": ref_e,
        "epol_rel_err_mixed": abs(mixed["epol"] - ref_e) / abs(ref_e) if ref_e else None,
        "mu_maxabs_err_mixed": mu_err(mixed["ipd"]),
        "roofline": {"bound": "mfma", "engine": "valu",
                     "compute": "compute-bound (the 'mfma' category of the two this line allows), executed on "
                                "the vector ALUs, not the matrix cores: rsqrt, exp and 3x3 tensor products per "
                                "pair; priced against the published FP32 peak, which counts packed FP32 "
                                "instructions that this build does not emit (vector_fma_measured gives the rate "
                                "non-packed FMAs reach on this device); FLOP = N(N-1)/2 (29 + 36m), the "
                                "pair-once count (each unordered pair once, m = 2 columns)",
                     "flop_formula": "N(N-1)/2*(29+36m)",
                     "kernel": "k_tmv_sym<float,2,AMOEBA> (on-the-fly FP32 pair kernel, each unordered pair "
                               "evaluated once)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS if achieved else None,
                     "traffic": profile_traffic("k_tmv_sym<float,2,AMOEBA>"),
                     "launches_timed": calls, "timed_every": PROF_EVERY, "avg_ms": avg_ms,
                     "flop_per_launch": flops,
                     "vector_fma_measured": vfma,
                     "frac_of_measured_fp32_vector_fma": achieved / vfma["fp32_tflops"] if achieved else None},
        "kernel_time_ms_timed": {k: v[1] for k, v in mixed["stats"].items()},
        "kernel_launches_timed": {k: v[0] for k, v in mixed["stats"].items()},
    }
    if multi is not None:
        out["multi_rhs"] = multi
    if large is not None:
        out["large_60k"] = large
    if serving is not None:
        out["serving"] = serving
    if grad is not None:
        out["geomgrad"] = grad
    if fmm is not None:
        out["fmm_250k"] = fmm
    if incore is not None:
        b = incore_bytes(n, 2, 4)
        out["incore_fp32_matvec"] = {
            "kernel": "k_tpack_mv<float,2> (packed symmetric T, 64x64 site tiles)", "bound": "hbm",
            "avg_ms_stream_kernel": incore[1], "avg_ms_with_finalize": incore[0], "bytes_per_launch": b,
            "achieved": b / (incore[1] * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": b / (incore[1] * 1e-3) / 1e9 / PEAK_HBM_GBS, "other_sizes": incore_extra}
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
            "fp64_pair_kernel": {"kernel": "k_tmv_sym<double,1,AMOEBA> (pair-once, one column per solve of "
                                           "the reference's D-then-P CG)",
                                 "launches_timed": c64, "timed_every": PROF_EVERY, "avg_ms": ms64 / max(c64, 1),
                                 "flop_per_launch": sym_pair_flops(n, 1), "flop_formula": "N(N-1)/2*(29+36m), m=1",
                                 "achieved_tflops": sym_pair_flops(n, 1) / (ms64 / max(c64, 1) * 1e-3) / 1e12
                                 if c64 else None, "peak": PEAK_FP64_TFLOPS,
                                 "frac": sym_pair_flops(n, 1) / (ms64 / max(c64, 1) * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
                                 if c64 else None,
                                 "frac_of_measured_fp64_vector_fma":
                                     sym_pair_flops(n, 1) / (ms64 / max(c64, 1) * 1e-3) / 1e12 / vfma["fp64_tflops"]
                                     if c64 else None},
        })
        if "cpu" in legs:
            print("[bench] cpu baseline", file=sys.stderr, flush=True)
            out["cpu_baseline"] = cpu_baseline(spec, args.tol)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
