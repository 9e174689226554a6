"""Golden vectors for the serving workload, from the oracle.

Workload: synth.serving_requests(16, seed=5, mean_gap_ms=10) -- the requests
bench.py's serving leg submits.  For the requests in PICK (the two largest,
39,855 and 35,721 sites, one of 8,719 and one of 633) the first and the last
solve of the SCF sequence are solved by the oracle's FP64 CG to 1e-12 with the
|p.Tp| quick return disabled, each from a zero guess.  Writes
tests/golden/serving16_seed5.json: per (request, step) E_pol and the dipoles of
a site sample.  About 35 minutes on 8 cores:

    python tests/golden/make_serving_golden.py
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

NREQ, SEED, GAP_MS = 16, 5, 10.0
PICK = (9, 13, 10, 4)


def main():
    sizes, nsteps, _, charges, make_spec = synth.serving_requests(NREQ, SEED, GAP_MS)
    oa.lib().oracle_set_small_step_eps(0.0)
    out = dict(nreq=NREQ, seed=SEED, mean_gap_ms=GAP_MS, tol=1e-12, cases=[])
    for k in PICK:
        spec = make_spec(k)
        o = oa.OracleSystem(synth.oracle_prep(spec))
        E = o.m2d()
        cpol = spec["cmm"][o.polar_mm]
        n = len(cpol)
        sample = sorted(set(list(range(0, 16)) + list(range(n // 2, n // 2 + 16)) + list(range(n - 16, n))))
        for step in (0, int(nsteps[k]) - 1):
            t = time.time()
            q = charges[k][step]
            ext = synth.point_charge_field(cpol, q[:, :3], q[:, 3])
            ipd = o.polarization(E + ext[None], use_guess=False, tol=1e-12, maxit=2000)
            out["cases"].append(dict(request=k, n=int(sizes[k]), step=step, epol=o.energy(ipd),
                                     iters=o.iters.tolist(), sample=sample,
                                     ipd_sample=ipd[:, sample, :].tolist()))
            print(k, sizes[k], step, out["cases"][-1]["epol"], o.iters, f"{time.time() - t:.1f}s", flush=True)
            with open(os.path.join(HERE, "serving16_seed5.json"), "w") as f:
                json.dump(out, f)


if __name__ == "__main__":
    main()
