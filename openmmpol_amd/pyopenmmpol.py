"""Python face of the library, mirroring the reference's pyopenmmpol module for
the polarization path (pythonapi/pyopenmmpol/pyommp_interface.cpp:9-30,
176-266, 482-530, 888-890, 1351-1536): same class, method and property names,
same argument meaning and the same ValueError checks.  Everything below calls
the C-ABI of libopenmmpol_amd.so.
"""
import ctypes

import numpy as np

from ._lib import c_dp, c_ip, lib

OMMP_SOLVER_NONE, OMMP_SOLVER_CG, OMMP_SOLVER_DIIS, OMMP_SOLVER_INVERSION = 0, 1, 2, 3
OMMP_MATV_NONE, OMMP_MATV_INCORE, OMMP_MATV_DIRECT = 0, 1, 2
OMMP_FF_AMOEBA, OMMP_FF_AMBER = 1, 0
PREC_FP64, PREC_MIXED = 0, 1

available_solvers = {"none": OMMP_SOLVER_NONE, "default": OMMP_SOLVER_CG, "conjugate gradient": OMMP_SOLVER_CG,
                     "cg": OMMP_SOLVER_CG, "inversion": OMMP_SOLVER_INVERSION, "diis": OMMP_SOLVER_DIIS}
available_matrix_vector = {"none": OMMP_MATV_NONE, "default": OMMP_MATV_DIRECT, "incore": OMMP_MATV_INCORE,
                           "direct": OMMP_MATV_DIRECT}
verbosity = {"none": 0, "low": 1, "high": 2, "debug": 3}
__version__ = "openmmpol-amd-0.1"


def set_verbose(level):
    lib().ommp_set_verbose(int(level))


def time_push():
    lib().ommp_time_push()


def time_pull(label):
    lib().ommp_time_pull(label.encode())


def set_outputfile(path):
    lib().ommp_set_outputfile(path.encode())


def close_outputfile():
    lib().ommp_close_outputfile()


def message(msg, level, prefix=""):
    lib().ommp_message(msg.encode(), int(level), prefix.encode())


def _view(ptr, shape):
    if not ptr:
        return np.zeros(shape)
    return np.ctypeslib.as_array(ptr, shape=shape)


class OMMPSystem:
    """System of OMMP library (polarization path)."""

    def __init__(self, mmp_filename=None, _handle=None):
        if _handle is not None:
            self.handler = _handle
        else:
            if mmp_filename is None:
                raise ValueError("mmp_filename is required")
            self.handler = lib().ommp_init_mmp(str(mmp_filename).encode())

    @classmethod
    def from_arrays(cls, amoeba, cmm, q, pol, bonds_ri, bonds_ci, pgroup=None, frames=None, thole_a=None,
                    ff_rules=0):
        """Build a system from arrays (ommp_amd_init_arrays).  cmm in bohr,
        q in the .mmp convention, pol in bohr^3, 1-2 connectivity as 0-based CSR."""
        nmm = len(pol)
        arrs = dict(cmm=np.ascontiguousarray(cmm, dtype=np.float64).reshape(nmm, 3),
                    q=np.ascontiguousarray(q, dtype=np.float64).reshape(nmm, -1),
                    pol=np.ascontiguousarray(pol, dtype=np.float64),
                    ri=np.ascontiguousarray(bonds_ri, dtype=np.int32),
                    ci=np.ascontiguousarray(bonds_ci if len(bonds_ci) else np.zeros(1), dtype=np.int32))
        pg = None if pgroup is None else np.ascontiguousarray(pgroup, dtype=np.int32)
        fr = None if frames is None else np.ascontiguousarray(frames, dtype=np.int32).reshape(nmm, 4)
        ta = None if thole_a is None else np.ascontiguousarray(thole_a, dtype=np.float64)
        h = lib().ommp_amd_init_arrays(OMMP_FF_AMOEBA if amoeba else OMMP_FF_AMBER, int(ff_rules), nmm,
                                       arrs["cmm"].ctypes.data_as(c_dp), arrs["q"].ctypes.data_as(c_dp),
                                       arrs["pol"].ctypes.data_as(c_dp),
                                       None if ta is None else ta.ctypes.data_as(c_dp),
                                       arrs["ri"].ctypes.data_as(c_ip), arrs["ci"].ctypes.data_as(c_ip),
                                       None if pg is None else pg.ctypes.data_as(c_ip),
                                       None if fr is None else fr.ctypes.data_as(c_ip))
        return cls(_handle=h)

    def __del__(self):
        self.terminate()

    def terminate(self):
        if getattr(self, "handler", None):
            lib().ommp_terminate(self.handler)
            self.handler = None

    @property
    def is_init(self):
        return bool(self.handler)

    # ------------------------------------------------------------- properties
    @property
    def pol_atoms(self):
        return lib().ommp_get_pol_atoms(self.handler)

    @property
    def mm_atoms(self):
        return lib().ommp_get_mm_atoms(self.handler)

    @property
    def n_ipd(self):
        return lib().ommp_get_n_ipd(self.handler)

    @property
    def _ld_cart(self):
        return lib().ommp_get_ld_cart(self.handler)

    @property
    def is_amoeba(self):
        return bool(lib().ommp_ff_is_amoeba(self.handler))

    @property
    def ipd(self):
        return _view(lib().ommp_get_ipd(self.handler), (self.n_ipd, self.pol_atoms, 3))

    @property
    def cmm(self):
        return _view(lib().ommp_get_cmm(self.handler), (self.mm_atoms, 3))

    @property
    def cpol(self):
        return _view(lib().ommp_get_cpol(self.handler), (self.pol_atoms, 3))

    @property
    def _q(self):
        return _view(lib().ommp_get_q(self.handler), (self.mm_atoms, self._ld_cart))

    @property
    def static_charges(self):
        return self._q[:, 0]

    @property
    def polar_mm(self):
        return _view(lib().ommp_get_polar_mm(self.handler), (self.pol_atoms,))

    @property
    def zmm(self):
        return _view(lib().ommp_get_zmm(self.handler), (self.mm_atoms,))

    # --------------------------------------------------------------- methods
    def set_external_field(self, external_field, nomm=False, solver="none", matv="none"):
        ef = np.ascontiguousarray(external_field, dtype=np.float64)
        if ef.ndim != 2 or ef.shape[0] != self.pol_atoms or ef.shape[1] != 3:
            raise ValueError("ext_field should be shaped [pol_atoms, 3]")
        if solver not in available_solvers:
            raise ValueError("Selected solver is not available!")
        if matv not in available_matrix_vector:
            raise ValueError("Selected matrix-vector method is not available!")
        fn = lib().ommp_set_external_field_nomm if nomm else lib().ommp_set_external_field
        fn(self.handler, ef.ctypes.data_as(c_dp), available_solvers[solver], available_matrix_vector[matv])

    def get_polelec_energy(self):
        return lib().ommp_get_polelec_energy(self.handler)

    def get_fixedelec_energy(self):
        """Interaction energy of the static multipoles (pyommp_interface.cpp:892-894)."""
        return lib().ommp_get_fixedelec_energy(self.handler)

    def get_full_ele_energy(self):
        """Fixed multipoles plus polarization (mod_interface.F90:388-398)."""
        return lib().ommp_get_full_ele_energy(self.handler)

    # potentials and fields at external points (pyommp_interface.cpp:362-480):
    # cext [n, 3] in bohr; new zeroed arrays are returned
    def _x2e(self, fn, cext, field):
        c = np.ascontiguousarray(cext, dtype=np.float64)
        if c.ndim != 2 or c.shape[1] != 3:
            raise ValueError("cext should be shaped [:, 3]")
        out = np.zeros((c.shape[0], 3) if field else c.shape[0])
        fn(self.handler, c.shape[0], c.ctypes.data_as(c_dp), out.ctypes.data_as(c_dp))
        return out

    def potential_mm2ext(self, cext):
        return self._x2e(lib().ommp_potential_mm2ext, cext, False)

    def potential_pol2ext(self, cext):
        return self._x2e(lib().ommp_potential_pol2ext, cext, False)

    def potential_mmpol2ext(self, cext):
        return self._x2e(lib().ommp_potential_mmpol2ext, cext, False)

    def field_mm2ext(self, cext):
        return self._x2e(lib().ommp_field_mm2ext, cext, True)

    def field_pol2ext(self, cext):
        return self._x2e(lib().ommp_field_pol2ext, cext, True)

    def field_mmpol2ext(self, cext):
        return self._x2e(lib().ommp_field_mmpol2ext, cext, True)

    # geometric gradients (pyommp_interface.cpp:518-606): [mm_atoms, 3], Hartree/bohr
    def _numerical_grad(self, energy, dd=1e-5):
        """Central differences of energy() over every MM coordinate, as the
        reference's numerical_grad; the coordinates are restored exactly."""
        c0 = np.array(self.cmm, dtype=np.float64)
        g = np.zeros_like(c0)
        try:
            for i in range(c0.shape[0]):
                for k in range(3):
                    c = c0.copy()
                    c[i, k] += dd
                    self.update_coordinates(c)
                    ep = energy()
                    c[i, k] = c0[i, k] - dd
                    self.update_coordinates(c)
                    g[i, k] = (ep - energy()) / (2.0 * dd)
        finally:
            self.update_coordinates(c0)
        return g

    def polelec_geomgrad(self, numerical=False):
        """d E_pol / d r of every MM atom (ommp_polelec_geomgrad)."""
        if numerical:
            return self._numerical_grad(self.get_polelec_energy)
        g = np.zeros((self.mm_atoms, 3))
        lib().ommp_polelec_geomgrad(self.handler, g.ctypes.data_as(c_dp))
        return g

    def rotation_geomgrad(self, E, Egrd):
        """Gradient through the multipole frames of the energy of the rotated
        multipoles in a field E [mm_atoms, 3] and field gradient Egrd [mm_atoms, 6]."""
        E = np.ascontiguousarray(E, dtype=np.float64)
        Egrd = np.ascontiguousarray(Egrd, dtype=np.float64)
        if E.shape != (self.mm_atoms, 3) or Egrd.shape != (self.mm_atoms, 6):
            raise ValueError("E should be shaped [mm_atoms, 3] and Egrd [mm_atoms, 6]")
        g = np.zeros((self.mm_atoms, 3))
        lib().ommp_rotation_geomgrad(self.handler, E.ctypes.data_as(c_dp), Egrd.ctypes.data_as(c_dp),
                                     g.ctypes.data_as(c_dp))
        return g

    def turn_pol_off(self, nopol):
        """Switch off the polarizability of the MM atoms nopol (1-based, as the
        reference's turn_pol_off); the polarizable set shrinks and the dipoles
        are dropped."""
        idx = np.ascontiguousarray(nopol, dtype=np.int32).reshape(-1)
        lib().ommp_turn_pol_off(self.handler, len(idx), idx.ctypes.data_as(c_ip))

    def update_coordinates(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        if c.ndim != 2 or c.shape[0] != self.mm_atoms or c.shape[1] != 3:
            raise ValueError("ext_field should be shaped [mm_atoms, 3]")
        lib().ommp_update_coordinates(self.handler, c.ctypes.data_as(c_dp))

    # fast multipoles (ommp_enable_fmm / ommp_set_fmm_*; lengths in bohr)
    def enable_fmm(self, lmax=None, lmax_pol=None, distance=None, min_cell_size=None):
        if lmax is not None:
            lib().ommp_set_fmm_lmax(self.handler, int(lmax))
        if lmax_pol is not None:
            lib().ommp_set_fmm_lmax_pol(self.handler, int(lmax_pol))
        if distance is not None:
            lib().ommp_set_fmm_distance(self.handler, float(distance))
        if min_cell_size is not None:
            lib().ommp_set_fmm_min_cell_size(self.handler, float(min_cell_size))
        lib().ommp_enable_fmm(self.handler)

    def disable_fmm(self):
        lib().ommp_disable_fmm(self.handler)

    @property
    def use_fmm(self):
        return bool(lib().ommp_use_fmm(self.handler))

    def set_default_solver(self, solver):
        lib().ommp_set_default_solver(self.handler, available_solvers[solver])

    def set_default_matv(self, matv):
        lib().ommp_set_default_matv(self.handler, available_matrix_vector[matv])

    # ------------------------------------------------------------ extensions
    def set_screening_parameters(self, m, p, d, u, intra=None):
        arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (m, p, d, u)]
        it = None if intra is None else np.ascontiguousarray(intra, dtype=np.float64)
        lib().ommp_amd_set_screening_parameters(self.handler, *[a.ctypes.data_as(c_dp) for a in arr],
                                                None if it is None else it.ctypes.data_as(c_dp))

    def set_solver_tolerance(self, tol, max_iter=200):
        lib().ommp_amd_set_solver_tolerance(self.handler, float(tol), int(max_iter))

    def set_inner_preconditioner(self, kind):
        """Inner FP32 solves of the mixed solver: "alpha" or "near" (default)."""
        lib().ommp_amd_set_inner_preconditioner(self.handler, {"alpha": 0, "near": 1}[kind])

    def inner_precond_apply(self, r):
        """z = M^-1 r of the near-field preconditioner (r: [ncols, pol_atoms, 3]);
        returns (z, in_use)."""
        x = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, self.pol_atoms, 3)
        z = np.zeros_like(x)
        on = lib().ommp_amd_inner_precond_apply(self.handler, x.shape[0], x.ctypes.data_as(c_dp),
                                                z.ctypes.data_as(c_dp))
        return z, bool(on)

    def set_precision(self, mode):
        lib().ommp_amd_set_precision(self.handler, {"fp64": PREC_FP64, "mixed": PREC_MIXED}[mode])

    @property
    def last_iterations(self):
        out = (ctypes.c_int32 * 4)()
        lib().ommp_amd_get_last_iterations(self.handler, out)
        return list(out)

    @property
    def last_status(self):
        return lib().ommp_amd_get_last_status(self.handler)

    @property
    def last_timings_ms(self):
        out = (ctypes.c_double * 3)()
        lib().ommp_amd_get_last_timings(self.handler, out)
        return list(out)

    @property
    def e_m2d(self):
        return _view(lib().ommp_amd_get_e_m2d(self.handler), (self.n_ipd, self.pol_atoms, 3)).copy()

    @property
    def thole(self):
        return _view(lib().ommp_amd_get_thole(self.handler), (self.mm_atoms,)).copy()

    @property
    def pol(self):
        return _view(lib().ommp_amd_get_pol(self.handler), (self.pol_atoms,)).copy()

    def get_list(self, which):
        """Host-side screening/topology lists (see ommp_amd_get_list)."""
        nnz = lib().ommp_amd_get_list(self.handler, int(which), None, None, None)
        nrow = self.pol_atoms if which == 2 else self.mm_atoms
        if which >= 7:
            # group lists: rows = groups; size from the CSR itself
            nrow = self.mm_atoms
        ri = np.zeros(nrow + 2, dtype=np.int32)
        ci = np.zeros(max(nnz, 1), dtype=np.int32)
        val = np.zeros(max(nnz, 1))
        lib().ommp_amd_get_list(self.handler, int(which), ri.ctypes.data_as(c_ip), ci.ctypes.data_as(c_ip),
                                val.ctypes.data_as(c_dp))
        return ri, ci[:nnz], val[:nnz]

    def tmatvec(self, x, kind="otf64", repeat=1):
        """y = T x for x [ncols, pol_atoms, 3]; returns (y, mean ms per launch)."""
        kinds = {"otf64": 0, "otf32": 1, "incore32": 2, "incore64": 3, "fmm": 4, "sym64": 5, "sym32": 6,
                 "plain64": 7, "plain32": 8}
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, self.pol_atoms, 3)
        y = np.zeros_like(x)
        ms = lib().ommp_amd_tmatvec(self.handler, kinds[kind], x.shape[0], x.ctypes.data_as(c_dp),
                                    y.ctypes.data_as(c_dp), int(repeat))
        return y, ms

    def block_tmatvec(self, x, repeat=1):
        """y = T x for up to 128 columns x [ncols, pol_atoms, 3] through the
        multi-RHS solver's matrix-core inner matvec; returns (y, mean ms per launch)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, self.pol_atoms, 3)
        if not 1 <= x.shape[0] <= 128:
            raise ValueError("block_tmatvec takes 1 to 128 columns")
        y = np.zeros_like(x)
        ms = lib().ommp_amd_block_tmatvec(self.handler, x.shape[0], x.ctypes.data_as(c_dp),
                                          y.ctypes.data_as(c_dp), int(repeat))
        return y, ms

    def profile_kernels(self, enable=True, every=1):
        """Bracket pair-kernel launches with HIP events (every `every`-th launch of
        each class when enabled)."""
        lib().ommp_amd_profile_kernels_every(self.handler, int(every) if enable else 0)

    def kernel_stats(self, reset=False):
        """{class: (launches, total ms)} for otf64, otf32, incore32, incore64, static_field,
        geomgrad."""
        out = (ctypes.c_double * 12)()
        lib().ommp_amd_get_kernel_stats(self.handler, out, bool(reset))
        names = ["otf64", "otf32", "incore32", "incore64", "static_field", "geomgrad"]
        return {nm: (int(out[2 * k]), out[2 * k + 1]) for k, nm in enumerate(names)}

    def set_external_field_device(self, d_ext_ptr, solver="none"):
        lib().ommp_amd_set_external_field_device(self.handler, ctypes.c_void_p(d_ext_ptr),
                                                 available_solvers[solver])

    def set_external_fields_multi(self, ext):
        """Solve nrhs external fields together (batched mixed solver).

        ext: [nrhs, pol_atoms, 3].  Returns (ipd [nrhs, n_ipd, pol_atoms, 3],
        epol [nrhs]).  The system's own dipoles are not changed."""
        ext = np.ascontiguousarray(ext, dtype=np.float64).reshape(-1, self.pol_atoms, 3)
        nrhs = ext.shape[0]
        ipd = np.zeros((nrhs, self.n_ipd, self.pol_atoms, 3))
        ep = np.zeros(nrhs)
        lib().ommp_amd_set_external_fields_multi(self.handler, nrhs, ext.ctypes.data_as(c_dp),
                                                 ipd.ctypes.data_as(c_dp), ep.ctypes.data_as(c_dp))
        return ipd, ep

    def set_external_fields_multi_device(self, nrhs, d_ext_ptr, d_ipd_ptr=None):
        """Device-pointer variant; returns the energies [nrhs]."""
        ep = np.zeros(nrhs)
        lib().ommp_amd_set_external_fields_multi_device(self.handler, int(nrhs), ctypes.c_void_p(d_ext_ptr),
                                                        ctypes.c_void_p(d_ipd_ptr or 0),
                                                        ep.ctypes.data_as(c_dp))
        return ep

    def charges_field(self, charges):
        """Field of point charges [nq, 4] (x, y, z, q) at the polarizable sites."""
        q = np.ascontiguousarray(charges, dtype=np.float64).reshape(-1, 4)
        out = np.zeros((self.pol_atoms, 3))
        lib().ommp_amd_charges_field(self.handler, q.shape[0], q.ctypes.data_as(c_dp), out.ctypes.data_as(c_dp))
        return out

    def device_ipd_ptr(self):
        return lib().ommp_amd_device_ipd(self.handler)

    def reset_guess(self, also_static_field=False):
        lib().ommp_amd_reset_guess(self.handler, bool(also_static_field))
