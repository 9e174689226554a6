"""Golden vectors for the multi-RHS workload, from the oracle.

System: 15,000 AMOEBA sites in a 33 A sphere with an empty 6 A central cavity
(seed 2); right-hand sides: E_M2D plus the field of a 30-charge QM region in
the cavity, displaced N(0, 0.1 A) per right-hand side
(synth.qm_charge_sets(64, seed=21)).  Writes tests/golden/multi15000_cavity.json
(E_pol and dipole samples of the right-hand sides in FIELDS -- the first four,
then every fourth up to the last, 19 of the 64 -- tight FP64 CG to 1e-12 with
the |p.Tp| quick return disabled) and multi15000_cavity_ipd0.npy (every dipole
of right-hand side 0).  About 15 minutes on 8 cores:

    python tests/golden/make_multi_golden.py
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

N, RADIUS, SEED, CAVITY, NRHS, QSEED = 15000, 33.0, 2, 6.0, 64, 21
FIELDS = [0, 1, 2, 3] + list(range(7, NRHS, 4))
SAMPLE = list(range(0, 48)) + list(range(7000, 7024)) + list(range(N - 24, N))


def main():
    spec = synth.sphere_system(N, RADIUS, seed=SEED, cavity=CAVITY)
    o = oa.OracleSystem(synth.oracle_prep(spec))
    oa.lib().oracle_set_small_step_eps(0.0)
    E = o.m2d()
    charges = synth.qm_charge_sets(NRHS, QSEED)
    cpol = spec["cmm"][o.polar_mm]
    out = dict(n=N, radius=RADIUS, seed=SEED, cavity=CAVITY, nrhs=NRHS, qm_seed=QSEED, tol=1e-12, sample=SAMPLE,
               rhs=[])
    for k in FIELDS:
        t = time.time()
        ext = synth.point_charge_field(cpol, charges[k, :, :3], charges[k, :, 3])
        ipd = o.polarization(E + ext[None], use_guess=False, tol=1e-12, maxit=1000)
        out["rhs"].append(dict(field=k, epol=o.energy(ipd), iters=o.iters.tolist(), ipd_sample=ipd[:, SAMPLE, :].tolist()))
        print(k, out["rhs"][-1]["epol"], o.iters, f"{time.time() - t:.1f}s", flush=True)
        if k == 0:
            np.save(os.path.join(HERE, "multi15000_cavity_ipd0.npy"), ipd)
    with open(os.path.join(HERE, "multi15000_cavity.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
