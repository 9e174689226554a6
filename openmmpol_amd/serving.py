"""Python face of the serving loop (ommp_amd_server_* in include/ommp_amd.h)."""
import ctypes

import numpy as np

from ._lib import c_dp, c_ip, lib


class Server:
    """Continuous-batching polarization server on the current GPU.

    submit(system, charges) queues an SCF-like request: charges is
    [nsteps, nq, 4] (x, y, z, q) and step k's external field is that of
    charges[k].  Requests run in two lanes by system size (below 4,096
    polarizable sites, and the rest); they are admitted in arrival order while
    their lane's `max_active` slots and the shared `hbm_budget` bytes allow, and
    the admitted solves of a lane share every launch of that lane."""

    def __init__(self, hbm_budget=64e9, max_active=16):
        self.h = lib().ommp_amd_server_create(float(hbm_budget), int(max_active))
        self._nsteps = {}
        self._keep = {}

    def submit(self, system, charges):
        q = np.ascontiguousarray(charges, dtype=np.float64)
        if q.ndim != 3 or q.shape[2] != 4:
            raise ValueError("charges must be [nsteps, nq, 4]")
        rid = lib().ommp_amd_server_submit(self.h, system.handler, q.shape[0], q.shape[1],
                                           q.ctypes.data_as(c_dp))
        self._nsteps[rid] = q.shape[0]
        self._keep[rid] = system
        return rid

    def wait(self, rid=-1):
        lib().ommp_amd_server_wait(self.h, int(rid))

    def result(self, rid):
        """(status, energies [nsteps], latency_ms [nsteps], iters [nsteps, 2])."""
        n = self._nsteps[rid]
        e = np.zeros(n)
        lat = np.zeros(n)
        it = np.zeros((n, 2), dtype=np.int32)
        st = lib().ommp_amd_server_result(self.h, int(rid), e.ctypes.data_as(c_dp), lat.ctypes.data_as(c_dp),
                                          it.ctypes.data_as(c_ip))
        return st, e, lat, it

    def stats(self):
        out = (ctypes.c_double * 6)()
        lib().ommp_amd_server_stats(self.h, out)
        keys = ["ticks", "solves", "mean_columns_per_tick", "max_admitted", "peak_bytes", "deferrals"]
        return dict(zip(keys, list(out)))

    def close(self):
        if self.h:
            lib().ommp_amd_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
