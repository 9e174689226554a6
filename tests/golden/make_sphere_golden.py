"""Golden vectors for the 15,000-site sphere system (seed 2), from the oracle.

Writes tests/golden/sphere15000_seed2.json (polarization energy and dipole
samples) and sphere15000_seed2_ipd_zero.npy (every dipole, zero external
field) of a tight FP64 CG solve (tolerance 1e-12 on the reference's
preconditioned rms, |p.Tp| quick return disabled), with zero external field
and with a 30-point-charge field.  Run on the CPU (about 3 minutes on 8 cores):

    python tests/golden/make_sphere_golden.py
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

N, RADIUS, SEED, QSEED = 15000, 33.0, 2, 11
SAMPLE = list(range(0, 64)) + list(range(7000, 7032)) + list(range(N - 32, N))


def main():
    spec = synth.sphere_system(N, RADIUS, seed=SEED)
    o = oa.OracleSystem(synth.oracle_prep(spec))
    oa.lib().oracle_set_small_step_eps(0.0)
    E = o.m2d()
    rng = np.random.default_rng(QSEED)
    cq, qq = synth.qm_charges(rng, 30, cavity=6.0)
    ext = synth.point_charge_field(spec["cmm"][o.polar_mm], cq, qq)
    out = dict(n=N, radius=RADIUS, seed=SEED, qm_seed=QSEED, tol=1e-12, sample=SAMPLE,
               e_m2d_sample=E[:, SAMPLE, :].tolist(), cases={})
    for label, field in (("zero", np.zeros_like(ext)), ("qm30", ext)):
        t = time.time()
        ipd = o.polarization(E + field[None], use_guess=False, tol=1e-12, maxit=1000)
        out["cases"][label] = dict(
            epol=o.energy(ipd), iters=o.iters.tolist(), ipd_sample=ipd[:, SAMPLE, :].tolist(),
            ipd_sum=[float(np.sum(ipd[k])) for k in range(2)],
            ipd_norm=[float(np.sqrt(np.sum(ipd[k] ** 2))) for k in range(2)])
        print(label, out["cases"][label]["epol"], o.iters, f"{time.time() - t:.1f}s", flush=True)
        if label == "zero":
            np.save(os.path.join(HERE, f"sphere{N}_seed{SEED}_ipd_zero.npy"), ipd)
    out["qm_charges"] = dict(pos_bohr=cq.tolist(), q=qq.tolist())
    with open(os.path.join(HERE, f"sphere{N}_seed{SEED}.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
