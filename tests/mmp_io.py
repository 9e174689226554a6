"""Test-side readers for the reference's fixture formats (numpy only).

This is test infrastructure: an independent Python restatement of the .mmp
reader (reference src/mod_inputloader.F90:15-299) and of the AMOEBA multipole
rotation (src/rotate_multipoles.F90:232-463), used to feed the oracle and to
cross-check the library's own C++ reader.
"""
import gzip
import os

import numpy as np

ANG2AU = 1.8897261245650  # config/openmmpol_const.h.in:34
AU2KCALMOL = 627.5096080306  # config/openmmpol_const.h.in:30
EPS_RP = np.finfo(np.float64).eps * 100.0  # src/mod_constants.F90:61

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REFERENCE_FIXTURES = os.path.join(GOLDEN, "reference")

# src/mod_constants.F90:37-52 (order: mscale, pscale, dscale, uscale, pscale_intra)
SCALES_AMOEBA = np.array([[0.0, 0.0, 0.4, 0.8], [0.0, 0.0, 1.0, 1.0], [0.0, 1.0, 1.0, 1.0],
                          [1.0, 1.0, 1.0, 1.0], [0.0, 0.0, 0.5, 1.0]])
SCALES_WANG_AL = np.array([[0.0, 0.0, 1.0, 1.0], [0.0, 0.0, 1.0, 1.0], [0.0, 0.0, 1.0, 1.0],
                           [0.0, 0.0, 1.0, 1.0], [1.0, 1.0, 1.0, 1.0]])
SCALES_WANG_DL = np.ones((5, 4))
THOLE_SCALE_WANG_AL = 2.5874
THOLE_SCALE_WANG_DL = 2.0580


def fixture_path(system, name):
    return os.path.join(REFERENCE_FIXTURES, system, name + ".gz")


def open_text(path):
    if path.endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path, "r")


class _ListDirected:
    """Fortran list-directed READ over a list of lines: every READ starts on a
    new record and continues over following lines until it has its values."""

    def __init__(self, lines):
        self.lines = lines
        self.pos = 0

    def skip(self, n):
        self.pos += n

    def read(self, k):
        vals = []
        while len(vals) < k:
            vals.extend(self.lines[self.pos].split())
            self.pos += 1
        return vals[:k]


def read_mmp(path):
    """Raw contents of a .mmp file (revisions 2 and 3)."""
    with open_text(path) as f:
        lines = f.read().splitlines()
    rd = _ListDirected(lines)
    rev = int(rd.read(1)[0])
    if rev not in (2, 3):
        raise ValueError("MMP file is only supported with version 2 or 3.")
    rd.skip(2)
    ff_type = int(rd.read(1)[0])
    ff_rules = int(rd.read(1)[0])
    rd.skip(17 if rev == 3 else 15)
    nmm = int(rd.read(1)[0])
    ld = 10 if ff_type == 1 else 1
    rd.skip(nmm + 1)
    cmm = np.array([[float(v) for v in rd.read(3)] for _ in range(nmm)])
    rd.skip(nmm)
    q = np.array([[float(v) for v in rd.read(ld)] for _ in range(nmm)])
    pol = np.array([float(rd.read(1)[0]) for _ in range(nmm)])
    i12 = np.array([[int(v) for v in rd.read(8)] for _ in range(nmm)], dtype=np.int64)
    out = dict(revision=rev, ff_type=ff_type, ff_rules=ff_rules, nmm=nmm, cmm=cmm, q=q,
               pol=pol, i12=i12)
    if ff_type == 1:
        out["ip11"] = np.array([[int(v) for v in rd.read(120)] for _ in range(nmm)], dtype=np.int64)
        out["frames"] = np.array([[int(v) for v in rd.read(4)] for _ in range(nmm)], dtype=np.int64)
    return out


def polgroup11_to_groups(ip11):
    """mod_inputloader.F90:259-299: 0-based group id per MM site."""
    nmm = ip11.shape[0]
    grp = np.full(nmm, -1, dtype=np.int64)
    g = 0
    for i in range(nmm):
        if grp[i] < 0:
            for m in ip11[i]:
                if m == 0:
                    break
                if grp[m - 1] >= 0:
                    raise ValueError("Unexpected error in polgroup11_to_mm2pg")
                grp[m - 1] = g
            g += 1
    if (grp < 0).any():
        raise ValueError("Unexpected error in polgroup11_to_mm2pg")
    return grp


def rotation_matrix(c, cx, cy, cz, frame):
    """rotation_matrix without derivatives (rotate_multipoles.F90:298-463)."""
    r = np.zeros((3, 3))
    if frame == 0:
        return np.eye(3)
    xi = cz - c
    eta = cx - c
    zeta = cy - c
    xn, en, zn = np.linalg.norm(xi), np.linalg.norm(eta), np.linalg.norm(zeta)
    u = np.zeros(3)
    v = np.zeros(3)
    if frame == 1:
        u, v = xi, eta
    elif frame == 2:
        u, v = en * xi + xn * eta, eta
    elif frame == 3:
        u = xi
        dot = u[2] / xn
        if dot <= 0.866 and abs(u[1]) > EPS_RP:
            v = np.array([1.0, 0.0, 0.0])
        else:
            v = np.array([0.0, 1.0, 0.0])
    elif frame == 4:
        u, v = xi, zn * eta + en * zeta
    elif frame == 5:
        u = en * zn * xi + xn * zn * eta + xn * en * zeta
        v = eta
    ez = u / np.sqrt(u @ u)
    ex = v - (v @ ez) * ez
    ex = ex / np.sqrt(ex @ ex)
    ey = np.array([ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2],
                   ez[0] * ex[1] - ez[1] * ex[0]])
    r[:, 0], r[:, 1], r[:, 2] = ex, ey, ez
    return r


def rotate_multipoles(cmm_au, q0, frames):
    """rotate_multipoles (rotate_multipoles.F90:232-295); q0 quadrupoles already /3."""
    q = q0.copy()
    n = cmm_au.shape[0]
    for j in range(n):
        fr, iz, ix, iy = frames[j]
        jz = iz - 1 if iz != 0 else j
        jx = ix - 1 if ix != 0 else j
        jy = iy - 1 if iy != 0 else j
        r = rotation_matrix(cmm_au[j], cmm_au[jx], cmm_au[jy], cmm_au[jz], fr)
        q[j, 1:4] = r @ q0[j, 1:4]
        Q = np.array([[q0[j, 4], q0[j, 5], q0[j, 7]],
                      [q0[j, 5], q0[j, 6], q0[j, 8]],
                      [q0[j, 7], q0[j, 8], q0[j, 9]]])
        R = r @ Q @ r.T
        q[j, 4], q[j, 6], q[j, 9] = R[0, 0], R[1, 1], R[2, 2]
        q[j, 5], q[j, 7], q[j, 8] = R[0, 1], R[0, 2], R[1, 2]
    return q


def adjacency_csr(i12):
    """adj_mat_from_conn (mod_adjacency_mat.F90:126-166), 0-based CSR, sorted rows."""
    rows = [sorted(int(v) - 1 for v in row if v != 0) for row in i12]
    ri = np.zeros(len(rows) + 1, dtype=np.int32)
    for i, r in enumerate(rows):
        ri[i + 1] = ri[i] + len(r)
    ci = np.array([v for r in rows for v in r], dtype=np.int32)
    return ri, ci


def prepare_mmp(path):
    """What ommp_init_mmp leaves in memory, as plain arrays for the oracle."""
    m = read_mmp(path)
    amoeba = m["ff_type"] == 1
    nmm = m["nmm"]
    cmm = m["cmm"] * ANG2AU
    raw_pol = m["pol"]
    pol = np.where(raw_pol > 1e-8, raw_pol, 0.0)  # mod_inputloader.F90:159-172
    if amoeba:
        pol = pol * ANG2AU ** 3
    ri, ci = adjacency_csr(m["i12"])
    out = dict(nmm=nmm, amoeba=amoeba, cmm=cmm, pol=pol, adj_ri=ri, adj_ci=ci,
               td=np.full(nmm, 0.39))
    if amoeba:
        q0 = m["q"].copy()
        q0[:, 4:10] /= 3.0  # mod_mmpol.F90:209-215
        out["q0"] = q0
        out["q"] = rotate_multipoles(cmm, q0, m["frames"])
        out["grp"] = polgroup11_to_groups(m["ip11"]).astype(np.int32)
        out["frames"] = m["frames"]
        out["scales"] = SCALES_AMOEBA.copy()
        out["thole_scale"] = 0.0
    else:
        out["q"] = m["q"].copy()
        out["grp"] = np.zeros(nmm, dtype=np.int32)
        if m["ff_rules"] == 1:
            out["scales"] = SCALES_WANG_DL.copy()
            out["thole_scale"] = THOLE_SCALE_WANG_DL
        else:
            out["scales"] = SCALES_WANG_AL.copy()
            out["thole_scale"] = THOLE_SCALE_WANG_AL
    out["polar_mm"] = np.nonzero(np.abs(pol) > EPS_RP)[0]
    return out


def read_ipd(path):
    """IPD_*.ref: one (x, y, z) row per polarizable site, D set then P set."""
    with open_text(path) as f:
        return np.array([[float(v) for v in line.split()] for line in f if line.strip()])


def read_ene(path):
    """ENE_*.ref: returns dict label -> value (EM, EP, ...)."""
    out = {}
    with open_text(path) as f:
        for line in f:
            if ":" in line:
                key = line.split()[0]
                out[key] = float(line.split(":")[-1])
    return out


def read_grad(path, term="EP"):
    """FULL_POTENTIAL.ref: the 'Grad <term>' block as an N x 3 array, in
    kcal/mol/angstrom (the energies above it are in Hartree)."""
    rows = None
    with open_text(path) as f:
        for line in f:
            if rows is None:
                t = line.split()
                if len(t) >= 2 and t[0] == "Grad" and t[1] == term:
                    rows = []
            elif line.strip():
                rows.append([float(v) for v in line.split()])
            else:
                break
    return None if rows is None else np.array(rows)


def read_ef(path):
    """EF_1.txt: field per MM atom (N x 3)."""
    with open_text(path) as f:
        return np.array([[float(v) for v in line.split()] for line in f if line.strip()])


def read_summary(path):
    """Parse the sections of a summary_*.ref file (format: mod_mmpol.F90:375-484)."""
    with open_text(path) as f:
        lines = f.read().splitlines()
    out = {"conn": {}, "pgconn": {}}
    i = 0
    def read_matrix(i):
        rows = []
        block = []
        i += 2  # section title, then the header with column numbers
        while i < len(lines):
            s = lines[i].strip()
            if not s:
                # blank separates column blocks
                if i + 1 < len(lines) and lines[i + 1].strip() and lines[i + 1].split()[0].isdigit() \
                        and "." not in lines[i + 1]:
                    rows.append(block)
                    block = []
                    i += 2
                    continue
                break
            if s[0].isalpha():
                break
            block.append([float(v) for v in s.split()])
            i += 1
        rows.append(block)
        mat = np.hstack([np.array(b) for b in rows if b])
        return mat, i
    while i < len(lines):
        line = lines[i]
        s = line.strip()
        if s.startswith(("mscale", "pscale", "dscale", "uscale")):
            key, vals = s.split(":")
            out[key.strip()] = np.array([float(v) for v in vals.split()])
            i += 1
        elif s.startswith("coordinates of polarizable"):
            out["cpol"], i = read_matrix(i)
        elif s.startswith("coordinates"):
            out["cmm"], i = read_matrix(i)
        elif s.startswith("multipoles - non rotated"):
            out["q0"], i = read_matrix(i)
        elif s.startswith("multipoles"):
            out["q"], i = read_matrix(i)
        elif s.startswith("polarizabilities"):
            out["pol"], i = read_matrix(i)
        elif s.startswith("thole factors"):
            out["thole"], i = read_matrix(i)
        elif s.startswith("mm_polar list") or s.startswith("polar_mm list"):
            key = s.split()[0]
            vals = []
            i += 1
            while i < len(lines) and lines[i].strip() and lines[i].strip()[0].isdigit():
                vals.extend(int(v) for v in lines[i].split())
                i += 1
            out[key] = np.array(vals)
        elif s.startswith("connectivity information for the"):
            atom = int(s.split()[4].split("-")[0]) - 1
            i += 1
            while i < len(lines):
                t = lines[i].strip()
                if t.startswith("1-") and "neighors" in t:
                    lvl = int(t.split()[0].split("-")[1])
                    kind = "pgconn" if "polarization" in t else "conn"
                    i += 1
                    vals = []
                    while i < len(lines) and "neighors" not in lines[i] and \
                            (not lines[i].strip() or lines[i].strip()[0].isdigit()):
                        if not lines[i].strip():
                            i += 1
                            break
                        vals.extend(int(v) - 1 for v in lines[i].split())
                        i += 1
                    out[kind].setdefault(atom, {})[lvl] = sorted(vals)
                else:
                    break
        else:
            i += 1
    return out
