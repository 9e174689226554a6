"""Golden vectors for the 60,000-site sphere (R 52 A, seed 3), from the oracle.

A tight FP64 CG solve (tolerance 1e-12 on the reference's preconditioned rms,
|p.Tp| quick return disabled) with zero external field.  Writes
tests/golden/sphere60000_seed3.json: E_pol, dipole samples, per-set sums and
norms of all dipoles, a static-field sample.  About 20 minutes on 8 cores:

    python tests/golden/make_sphere60k_golden.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))  # tests/
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_api as oa  # noqa: E402
from openmmpol_amd import synth  # noqa: E402

N, RADIUS, SEED = 60000, 52.0, 3
SAMPLE = list(range(0, 64)) + list(range(30000, 30032)) + list(range(N - 32, N))


def main():
    t = time.time()
    spec = synth.sphere_system(N, RADIUS, seed=SEED)
    o = oa.OracleSystem(synth.oracle_prep(spec))
    oa.lib().oracle_set_small_step_eps(0.0)
    E = o.m2d()
    print("static field", f"{time.time() - t:.1f}s", flush=True)
    t = time.time()
    ipd = o.polarization(E, use_guess=False, tol=1e-12, maxit=1000)
    out = dict(n=N, radius=RADIUS, seed=SEED, tol=1e-12, sample=SAMPLE, e_m2d_sample=E[:, SAMPLE, :].tolist(),
               epol=o.energy(ipd), iters=o.iters.tolist(), ipd_sample=ipd[:, SAMPLE, :].tolist(),
               ipd_sum=[float(np.sum(ipd[k])) for k in range(2)],
               ipd_norm=[float(np.sqrt(np.sum(ipd[k] ** 2))) for k in range(2)])
    print("zero", out["epol"], o.iters, f"{time.time() - t:.1f}s", flush=True)
    with open(os.path.join(HERE, f"sphere{N}_seed{SEED}.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
