"""ctypes wrapper over oracle/liboracle.so (CPU FP64 restatement of the reference).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the reference results are compared with, never as the thing
measured.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a, typ):
    return a.ctypes.data_as(typ)


def build():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double,
                                    _ip, _ip, _ip, _dp]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_npol.argtypes = [ctypes.c_void_p]
        L.oracle_turn_pol_off.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.oracle_turn_pol_off.restype = ctypes.c_int
        L.oracle_fatal_flag.argtypes = [ctypes.c_void_p]
        L.oracle_fatal_msg.argtypes = [ctypes.c_void_p]
        L.oracle_fatal_msg.restype = ctypes.c_char_p
        L.oracle_get_polar_mm.argtypes = [ctypes.c_void_p, _ip]
        L.oracle_get_thole.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_list_nnz.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_list_get.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip, _dp]
        L.oracle_screening_rule.restype = ctypes.c_double
        L.oracle_screening_rule.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int]
        L.oracle_damped_kernel.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_multipole_potential_field.restype = ctypes.c_double
        L.oracle_multipole_potential_field.argtypes = [_dp, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_m2d.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_tmatvec_otf.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.oracle_create_tmat.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_polarization.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip]
        L.oracle_energy_mm.restype = ctypes.c_double
        L.oracle_energy_mm.argtypes = [ctypes.c_void_p]
        L.oracle_energy_pol.restype = ctypes.c_double
        L.oracle_energy_pol.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_charges_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_set_small_step_eps.argtypes = [ctypes.c_double]
        L.oracle_x2e.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.oracle_geomgrad.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.oracle_rotation_geomgrad.restype = ctypes.c_int
        L.oracle_rotation_geomgrad.argtypes = [ctypes.c_int, _dp, _dp, _ip, _dp, _dp, _dp]
        _lib = L
    return _lib


SOLVER_CG, SOLVER_DIIS, SOLVER_INVERSION = 1, 2, 3
MATV_INCORE, MATV_DIRECT = 1, 2


class OracleSystem:
    """Prepared system in the oracle plus the reference's API-level state machine
    (ommp_set_external_field / ommp_get_polelec_energy, mod_interface.F90:184-370)."""

    def __init__(self, prep):
        L = lib()
        self.prep = prep
        self.amoeba = bool(prep["amoeba"])
        self.nmm = int(prep["nmm"])
        self._arrays = dict(
            cmm=np.ascontiguousarray(prep["cmm"], dtype=np.float64),
            q=np.ascontiguousarray(prep["q"], dtype=np.float64),
            pol=np.ascontiguousarray(prep["pol"], dtype=np.float64),
            td=np.ascontiguousarray(prep.get("td", np.full(self.nmm, 0.39)), dtype=np.float64),
            ri=np.ascontiguousarray(prep["adj_ri"], dtype=np.int32),
            ci=np.ascontiguousarray(prep["adj_ci"] if len(prep["adj_ci"]) else np.zeros(1), dtype=np.int32),
            grp=np.ascontiguousarray(prep["grp"], dtype=np.int32),
            scales=np.ascontiguousarray(prep["scales"], dtype=np.float64),
        )
        a = self._arrays
        self.h = L.oracle_create(self.nmm, int(self.amoeba), _ptr(a["cmm"], _dp), _ptr(a["q"], _dp),
                                 _ptr(a["pol"], _dp), _ptr(a["td"], _dp), float(prep.get("thole_scale", 0.0)),
                                 _ptr(a["ri"], _ip), _ptr(a["ci"], _ip), _ptr(a["grp"], _ip),
                                 _ptr(a["scales"], _dp))
        self.npol = L.oracle_npol(self.h)
        self.n_ipd = 2 if self.amoeba else 1
        self.ipd = np.zeros((self.n_ipd, self.npol, 3))
        self.ipd_done = False
        self.ipd_use_guess = False
        self.e_m2d = None
        self.iters = None
        self.tol = 1e-8
        self.maxit = 200

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def turn_pol_off(self, nopol):
        """ommp_turn_pol_off on the listed MM atoms (1-based), see oracle.c."""
        idx = np.ascontiguousarray(nopol, dtype=np.int32)
        lib().oracle_turn_pol_off(self.h, len(idx), _ptr(idx, _ip) if len(idx) else None)
        self.check()
        self.npol = lib().oracle_npol(self.h)
        self.ipd = np.zeros((self.n_ipd, self.npol, 3))
        self.ipd_done = False
        self.ipd_use_guess = False
        self.e_m2d = None

    def check(self):
        if lib().oracle_fatal_flag(self.h):
            raise RuntimeError(lib().oracle_fatal_msg(self.h).decode())

    @property
    def polar_mm(self):
        out = np.zeros(self.npol, dtype=np.int32)
        lib().oracle_get_polar_mm(self.h, _ptr(out, _ip))
        return out

    def thole(self):
        out = np.zeros(self.nmm)
        lib().oracle_get_thole(self.h, _ptr(out, _dp))
        return out

    def list_csr(self, which):
        """0=S_P_P 1=S_P_D 2=P_P, 3..6 = conn 1-2..1-5, 7..10 = group distance 0..3."""
        nnz = lib().oracle_list_nnz(self.h, which)
        nrow = self.npol if which == 2 else (self.nmm if which < 7 or which == 11
                                             else int(self.prep["grp"].max()) + 1)
        ri = np.zeros(nrow + 1, dtype=np.int32)
        ci = np.zeros(max(nnz, 1), dtype=np.int32)
        val = np.zeros(max(nnz, 1))
        lib().oracle_list_get(self.h, which, _ptr(ri, _ip), _ptr(ci, _ip), _ptr(val, _dp))
        return ri, ci[:nnz], val[:nnz]

    def screening_rule(self, i, kind_i, j, kind_j, field):
        return lib().oracle_screening_rule(self.h, i, ord(kind_i), j, ord(kind_j), ord(field))

    def damped_kernel(self, i_mm, j_mm, maxder):
        res = np.zeros(maxder + 1)
        dr = np.zeros(3)
        lib().oracle_damped_kernel(self.h, i_mm, j_mm, maxder, _ptr(res, _dp), _ptr(dr, _dp))
        return res, dr

    def m2d(self):
        E = np.zeros((self.n_ipd, self.npol, 3))
        lib().oracle_m2d(self.h, _ptr(E, _dp))
        self.check()
        return E

    def prepare_polelec(self):
        if self.e_m2d is None:
            self.e_m2d = self.m2d()
        return self.e_m2d

    def tmatvec(self, x, dodiag=True):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.npol, 3)
        y = np.zeros_like(x)
        lib().oracle_tmatvec_otf(self.h, _ptr(x, _dp), _ptr(y, _dp), int(dodiag))
        self.check()
        return y

    def tmat(self):
        n = 3 * self.npol
        T = np.zeros((n, n))
        lib().oracle_create_tmat(self.h, _ptr(T, _dp))
        self.check()
        return T

    def polarization(self, e, solver=SOLVER_CG, matv=MATV_DIRECT, use_guess=None, mask=3,
                     tol=None, maxit=None):
        e = np.ascontiguousarray(e, dtype=np.float64).reshape(self.n_ipd, self.npol, 3)
        if use_guess is None:
            use_guess = self.ipd_use_guess
        ipd = self.ipd.copy()
        iters = np.zeros(2, dtype=np.int32)
        rc = lib().oracle_polarization(self.h, _ptr(e, _dp), _ptr(ipd, _dp), solver, matv,
                                       int(use_guess), mask, self.tol if tol is None else tol,
                                       self.maxit if maxit is None else maxit, _ptr(iters, _ip))
        self.check()
        if rc:
            raise RuntimeError("Iterative solver did not converged")
        self.ipd = ipd
        self.iters = iters[: self.n_ipd].copy()
        self.ipd_done = True
        self.ipd_use_guess = True
        return ipd

    def set_external_field(self, ext, solver=0, matv=0, nomm=False, tol=None):
        """ommp_set_external_field[_nomm] (mod_interface.F90:184-233)."""
        ext = np.asarray(ext, dtype=np.float64).reshape(self.npol, 3)
        self.ipd_done = False
        if not nomm:
            e = self.prepare_polelec() + ext[None, :, :]
            # the C entry point never forwards matv (mod_interface.F90:221)
            return self.polarization(e, solver=solver or SOLVER_CG, matv=MATV_DIRECT, tol=tol)
        e = np.zeros((self.n_ipd, self.npol, 3))
        e[0] = ext
        return self.polarization(e, solver=solver or SOLVER_CG, matv=matv or MATV_DIRECT, mask=1, tol=tol)

    def get_polelec_energy(self):
        """ommp_get_polelec_energy (mod_interface.F90:343-370)."""
        if self.npol == 0:
            return 0.0
        if not self.ipd_done:
            self.polarization(self.prepare_polelec())
        ipd = np.ascontiguousarray(self.ipd)
        em = np.ascontiguousarray(self.prepare_polelec())
        return lib().oracle_energy_pol(self.h, _ptr(ipd, _dp), _ptr(em, _dp))

    def energy_mm(self):
        """energy_MM_MM: interaction energy of the static multipoles (Hartree)."""
        e = lib().oracle_energy_mm(self.h)
        self.check()
        return e

    def energy(self, ipd):
        ipd = np.ascontiguousarray(ipd, dtype=np.float64)
        em = np.ascontiguousarray(self.prepare_polelec())
        return lib().oracle_energy_pol(self.h, _ptr(ipd, _dp), _ptr(em, _dp))

    def x2e(self, cext, which, ipd=None):
        """Potential [n] and field [n, 3] at external points (potential_M2E /
        D2E, field_M2E / D2E): which 1 = multipoles, 2 = induced dipoles, 3 = both."""
        cext = np.ascontiguousarray(cext, dtype=np.float64).reshape(-1, 3)
        n = cext.shape[0]
        V = np.zeros(n)
        E = np.zeros((n, 3))
        ip = np.ascontiguousarray(self.ipd if ipd is None else ipd, dtype=np.float64)
        lib().oracle_x2e(self.h, _ptr(ip, _dp), n, _ptr(cext, _dp), int(which), _ptr(V, _dp), _ptr(E, _dp))
        self.check()
        return V, E

    def geomgrad_parts(self, ipd=None):
        """Translational part of polelec_geomgrad [nmm, 3] plus the field E_D2M
        [nmm, 3] and field gradient Egrd_D2M [nmm, 6] the rotation term needs."""
        ip = np.ascontiguousarray(self.ipd if ipd is None else ipd, dtype=np.float64)
        g = np.zeros((self.nmm, 3))
        E = np.zeros((self.nmm, 3))
        G = np.zeros((self.nmm, 6))
        lib().oracle_geomgrad(self.h, _ptr(ip, _dp), _ptr(g, _dp), _ptr(E, _dp), _ptr(G, _dp))
        self.check()
        return g, E, G

    def polelec_geomgrad(self, ipd=None):
        """polelec_geomgrad (mod_geomgrad.F90:113-273): dE_pol/dr [nmm, 3] in
        Hartree/bohr, rotation term included when the prep carries frames."""
        g, E, G = self.geomgrad_parts(ipd)
        if self.amoeba and "frames" in self.prep:
            g += rotation_geomgrad(self.prep["cmm"], self.prep["q0"], self.prep["frames"], E, G)
        return g

    def charges_field(self, cq, qq):
        cq = np.ascontiguousarray(cq, dtype=np.float64)
        qq = np.ascontiguousarray(qq, dtype=np.float64)
        E = np.zeros((self.npol, 3))
        lib().oracle_charges_field(self.h, len(qq), _ptr(cq, _dp), _ptr(qq, _dp), _ptr(E, _dp))
        return E


def frames_0based(frames):
    """.mmp frame rows (convention, z, x, y; 1-based, 0 = none) -> 0-based, -1 = none."""
    f = np.asarray(frames, dtype=np.int64).reshape(-1, 4)
    out = np.empty_like(f, dtype=np.int32)
    out[:, 0] = f[:, 0]
    out[:, 1:] = f[:, 1:] - 1
    return np.ascontiguousarray(out)


def rotation_geomgrad(cmm, q0, frames, E, Egrd):
    """rotation_geomgrad (rotate_multipoles.F90:3-230): [nmm, 3] gradient of the
    rotated multipoles' energy in the field E [nmm, 3] and gradient Egrd [nmm, 6]."""
    cmm = np.ascontiguousarray(cmm, dtype=np.float64)
    q0 = np.ascontiguousarray(q0, dtype=np.float64)
    fr = frames_0based(frames)
    E = np.ascontiguousarray(E, dtype=np.float64)
    Egrd = np.ascontiguousarray(Egrd, dtype=np.float64)
    g = np.zeros((cmm.shape[0], 3))
    rc = lib().oracle_rotation_geomgrad(cmm.shape[0], _ptr(cmm, _dp), _ptr(q0, _dp), _ptr(fr, _ip),
                                        _ptr(E, _dp), _ptr(Egrd, _dp), _ptr(g, _dp))
    if rc:
        raise RuntimeError("Ill defined rotation axis (either u or v vectors have 0 norm.")
    return g
