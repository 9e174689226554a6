"""Seeded synthetic systems for the polarization path.

* water_box: rigid waters on a jittered cubic lattice (non-periodic), one
  polarization group per water, AMOEBA default scale factors.
* sphere_system: N sites uniform in a sphere with a minimum pair distance, a
  nearest-neighbour chain graph with short branches (degree 1..4) giving the
  1-2..1-5 lists, polarization groups of 3..12 connected sites, random
  multipoles and polarizabilities, an optional empty cavity at the centre.
* point_charge_field: field of a set of point charges at the polarizable
  sites, undamped and unscaled, as the QM-nuclei field E_n2p of the reference
  (src/mod_qm_helper.F90:376-391).

All randomness comes from numpy.random.default_rng(seed); reductions that
feed the system use math.fsum so that the same seed gives the same bits on
every machine.  Lengths are returned in bohr, polarizabilities in bohr^3,
multipoles in the .mmp convention (quadrupoles x3 of the stored ones).
"""
import math

import numpy as np
from scipy.spatial import cKDTree

ANG2AU = 1.8897261245650
POL_CHOICES_A3 = (0.34, 0.496, 1.073, 1.334, 2.8)


def _random_rotations(rng, n):
    # uniform rotations from unit quaternions
    q = rng.normal(size=(n, 4))
    q /= np.sqrt((q * q).sum(axis=1))[:, None]
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], axis=-1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], axis=-1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=-1),
    ], axis=1)


def _traceless(rng, n, sigma):
    a = rng.normal(0.0, sigma, size=(n, 3, 3))
    a = 0.5 * (a + a.transpose(0, 2, 1))
    tr = (a[:, 0, 0] + a[:, 1, 1] + a[:, 2, 2]) / 3.0
    for k in range(3):
        a[:, k, k] -= tr
    # storage xx, xy, yy, xz, yz, zz
    return np.stack([a[:, 0, 0], a[:, 0, 1], a[:, 1, 1], a[:, 0, 2], a[:, 1, 2], a[:, 2, 2]], axis=1)


def _csr(nb):
    ri = np.zeros(len(nb) + 1, dtype=np.int32)
    for i, r in enumerate(nb):
        ri[i + 1] = ri[i] + len(r)
    ci = np.array([v for r in nb for v in sorted(r)], dtype=np.int32)
    return ri, ci


def water_box(nside=6, spacing=3.1, jitter=0.3, seed=1, dip_sigma=0.02, quad_sigma=0.05):
    """216 waters (648 sites) for nside=6.  Returns a dict for OMMPSystem.from_arrays."""
    rng = np.random.default_rng(seed)
    nw = nside ** 3
    g = np.arange(nside) * spacing
    centers = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    centers = centers + rng.uniform(-jitter, jitter, size=(nw, 3))
    roh, ang = 0.957, np.deg2rad(104.5)
    local = np.array([[0.0, 0.0, 0.0],
                      [roh * np.sin(ang / 2), 0.0, roh * np.cos(ang / 2)],
                      [-roh * np.sin(ang / 2), 0.0, roh * np.cos(ang / 2)]])
    R = _random_rotations(rng, nw)
    xyz = (centers[:, None, :] + np.einsum("wij,aj->wai", R, local)).reshape(-1, 3)
    n = 3 * nw
    q = np.zeros((n, 10))
    q[0::3, 0] = -0.51966
    q[1::3, 0] = 0.25983
    q[2::3, 0] = 0.25983
    q[:, 1:4] = rng.normal(0.0, dip_sigma, size=(n, 3))
    q[:, 4:10] = 3.0 * _traceless(rng, n, quad_sigma)
    pol = np.tile(np.array([0.837, 0.496, 0.496]), nw) * ANG2AU ** 3
    nb = [[] for _ in range(n)]
    for w in range(nw):
        o = 3 * w
        for h in (o + 1, o + 2):
            nb[o].append(h)
            nb[h].append(o)
    ri, ci = _csr(nb)
    grp = np.repeat(np.arange(nw, dtype=np.int32), 3)
    return dict(amoeba=True, cmm=xyz * ANG2AU, q=q, pol=pol, bonds_ri=ri, bonds_ci=ci, pgroup=grp,
                thole_a=np.full(n, 0.39), name=f"water{nw}", seed=seed)


def _sample_sphere(rng, n, radius, dmin, cavity):
    pts = np.zeros((0, 3))
    cell = dmin
    grid = {}
    out = []
    batch = 4096
    while len(out) < n:
        cand = rng.uniform(-radius, radius, size=(batch, 3))
        r2 = cand[:, 0] * cand[:, 0] + cand[:, 1] * cand[:, 1] + cand[:, 2] * cand[:, 2]
        ok = (r2 <= radius * radius) & (r2 >= cavity * cavity)
        for p in cand[ok]:
            key = tuple(np.floor(p / cell).astype(np.int64))
            clash = False
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for qv in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                            d = p - qv
                            if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < dmin * dmin:
                                clash = True
                                break
                        if clash:
                            break
                    if clash:
                        break
                if clash:
                    break
            if not clash:
                grid.setdefault(key, []).append(p)
                out.append(p)
                if len(out) == n:
                    break
    return np.array(out)


def sphere_system(n=15000, radius=33.0, dmin=1.0, seed=2, cavity=0.0, pscale=(0.0, 0.0, 0.4, 0.8),
                  q_sigma=0.3, dip_sigma=0.05, quad_sigma=0.1, bond_cut=2.2, branch_p=0.2):
    """Sites in a sphere (Angstrom inputs), chain topology, groups of 3..12 sites."""
    rng = np.random.default_rng(seed)
    pts = _sample_sphere(rng, n, radius, dmin, cavity)
    tree = cKDTree(pts)
    # chains: walk to the nearest unvisited site within bond_cut, from both ends
    visited = np.zeros(n, dtype=bool)
    order = []
    chains = []
    nb = [set() for _ in range(n)]
    starts = rng.permutation(n)

    def nearest_free(cur):
        best, bd = -1, 1e30
        for c in tree.query_ball_point(pts[cur], bond_cut):
            if visited[c]:
                continue
            d = pts[c] - pts[cur]
            d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            if d2 < bd or (d2 == bd and c < best):
                best, bd = c, d2
        return best

    for s in starts:
        if visited[s]:
            continue
        chain = [s]
        visited[s] = True
        maxlen = int(rng.integers(20, 200))
        for end in (1, 0):  # grow the tail, then the head
            while len(chain) < maxlen:
                cur = chain[-1] if end else chain[0]
                best = nearest_free(cur)
                if best < 0:
                    break
                nb[cur].add(best)
                nb[best].add(cur)
                visited[best] = True
                if end:
                    chain.append(best)
                else:
                    chain.insert(0, best)
        chains.append(chain)
        order.extend(chain)
    # renumber in chain order (spatially coherent indices)
    perm = np.array(order)
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    pts = pts[perm]
    nb = [set(int(inv[v]) for v in nb[int(perm[i])]) for i in range(n)]
    chains = [[int(inv[v]) for v in c] for c in chains]
    tree = cKDTree(pts)
    # branches and attachment of isolated sites (degree <= 4)
    for i in range(n):
        if len(nb[i]) >= 4:
            continue
        if len(nb[i]) > 0 and rng.random() >= branch_p:
            continue
        cand = tree.query_ball_point(pts[i], bond_cut if nb[i] else 2.0 * bond_cut)
        best, bd = -1, 1e30
        for c in cand:
            if c == i or c in nb[i] or len(nb[c]) >= 4:
                continue
            d = pts[c] - pts[i]
            d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            if d2 < bd or (d2 == bd and c < best):
                best, bd = c, d2
        if best >= 0:
            nb[i].add(best)
            nb[best].add(i)
    # polarization groups: consecutive chain segments of 3..10 sites; chains
    # shorter than 3 join the group of a bonded site outside them (<= 12)
    grp = np.full(n, -1, dtype=np.int64)
    g = 0
    small = []
    for chain in chains:
        L = len(chain)
        if L < 3:
            small.append(chain)
            continue
        k = 0
        while k < L:
            rem = L - k
            if rem <= 10:
                size = rem
            else:
                size = int(rng.integers(3, 11))
                if rem - size < 3:
                    size = rem - 3
            for v in chain[k:k + size]:
                grp[v] = g
            g += 1
            k += size
    counts = list(np.bincount(grp[grp >= 0], minlength=g)) if g else []
    for chain in small:
        cset = set(chain)
        target = -1
        for v in chain:
            for w in sorted(nb[v]):
                if w not in cset and grp[w] >= 0 and counts[grp[w]] + len(chain) <= 12:
                    target = grp[w]
                    break
            if target >= 0:
                break
        if target < 0:
            # bond to the nearest free-degree site of a group with room
            v = chain[0]
            dist, idx = tree.query(pts[v], k=min(n, 64))
            for c in idx:
                c = int(c)
                if c in cset or grp[c] < 0 or len(nb[c]) >= 4 or len(nb[v]) >= 4:
                    continue
                if counts[grp[c]] + len(chain) <= 12:
                    nb[v].add(c)
                    nb[c].add(v)
                    target = grp[c]
                    break
        if target < 0:
            target = g
            g += 1
            counts.append(0)
        for v in chain:
            grp[v] = target
        counts[target] += len(chain)
    nbl = [sorted(s) for s in nb]
    ri, ci = _csr(nbl)
    qq = rng.normal(0.0, q_sigma, size=n)
    qq = qq - math.fsum(qq) / n
    q = np.zeros((n, 10))
    q[:, 0] = qq
    q[:, 1:4] = rng.normal(0.0, dip_sigma, size=(n, 3))
    q[:, 4:10] = 3.0 * _traceless(rng, n, quad_sigma)
    pol = np.array(POL_CHOICES_A3)[rng.integers(0, len(POL_CHOICES_A3), size=n)] * ANG2AU ** 3
    scales = dict(m=[0.0, 0.0, 0.4, 0.8], p=list(pscale), d=[0.0, 1.0, 1.0, 1.0], u=[1.0, 1.0, 1.0, 1.0],
                  intra=[0.0, 0.0, 0.5, 1.0])
    return dict(amoeba=True, cmm=pts * ANG2AU, q=q, pol=pol, bonds_ri=ri, bonds_ci=ci,
                pgroup=grp.astype(np.int32), thole_a=np.full(n, 0.39), scales=scales,
                name=f"sphere{n}", seed=seed, radius=radius, cavity=cavity)


def qm_charges(rng, nq=30, cavity=6.0, sigma_q=0.5):
    """nq point charges N(0, sigma_q), net zero, inside a cavity (Angstrom radius)."""
    pos = []
    while len(pos) < nq:
        p = rng.uniform(-cavity, cavity, size=3)
        if p[0] * p[0] + p[1] * p[1] + p[2] * p[2] <= (0.8 * cavity) ** 2:
            pos.append(p)
    qq = rng.normal(0.0, sigma_q, size=nq)
    qq = qq - math.fsum(qq) / nq
    return np.array(pos) * ANG2AU, qq


def qm_charge_sets(nrhs, seed, nq=30, cavity=6.0, sigma_q=0.5, sigma_disp=0.1):
    """A 'QM region' of nq charges in the central cavity and nrhs copies of it:
    copy 0 as drawn, copy k > 0 with every charge displaced N(0, sigma_disp)
    Angstrom.  Returns [nrhs, nq, 4] (x, y, z in bohr, q in e)."""
    rng = np.random.default_rng(seed)
    cq, qq = qm_charges(rng, nq, cavity=cavity, sigma_q=sigma_q)
    out = np.zeros((nrhs, nq, 4))
    for k in range(nrhs):
        d = rng.normal(0.0, sigma_disp, size=(nq, 3)) * ANG2AU if k > 0 else 0.0
        out[k, :, :3] = cq + d
        out[k, :, 3] = qq
    return out


def point_charge_field(cpol_au, cq_au, qq):
    """E_j = sum_c q_c (r_j - r_c) / |r_j - r_c|^3 at every polarizable site (a.u.)."""
    E = np.zeros_like(cpol_au)
    for c in range(len(qq)):
        d = cpol_au - cq_au[c]
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        E += (qq[c] / (r2 * np.sqrt(r2)))[:, None] * d
    return E


def make_system(spec):
    """OMMPSystem from a generator dict (applies non-default scale factors)."""
    from .pyopenmmpol import OMMPSystem
    s = OMMPSystem.from_arrays(spec["amoeba"], spec["cmm"], spec["q"], spec["pol"], spec["bonds_ri"],
                               spec["bonds_ci"], pgroup=spec.get("pgroup"), thole_a=spec.get("thole_a"))
    sc = spec.get("scales")
    if sc is not None:
        s.set_screening_parameters(sc["m"], sc["p"], sc["d"], sc["u"], sc["intra"] if spec["amoeba"] else None)
    return s


def oracle_prep(spec):
    """The same system as plain arrays for the oracle (quadrupoles divided by 3)."""
    q = spec["q"].copy()
    if spec["amoeba"]:
        q[:, 4:10] /= 3.0
    sc = spec.get("scales")
    if sc is None:
        scales = np.array([[0.0, 0.0, 0.4, 0.8], [0.0, 0.0, 1.0, 1.0], [0.0, 1.0, 1.0, 1.0],
                           [1.0, 1.0, 1.0, 1.0], [0.0, 0.0, 0.5, 1.0]])
    else:
        scales = np.array([sc["m"], sc["p"], sc["d"], sc["u"], sc["intra"]], dtype=float)
    n = len(spec["pol"])
    return dict(nmm=n, amoeba=spec["amoeba"], cmm=spec["cmm"], q=q, pol=spec["pol"],
                adj_ri=spec["bonds_ri"], adj_ci=spec["bonds_ci"],
                grp=spec["pgroup"] if spec.get("pgroup") is not None else np.zeros(n, dtype=np.int32),
                td=spec.get("thole_a", np.full(n, 0.39)), scales=scales, thole_scale=0.0)


def serving_requests(nreq=16, seed=5, mean_gap_ms=10.0):
    """The serving workload: nreq SCF-like requests on spheres of log-uniform
    size 500-40,000 sites (density 0.1 / A^3, 6 A cavity, system seed 1000 + k),
    10-20 solves each, the field of step j that of a 30-charge QM region
    (qm_charge_sets(1, 2000 + k)) with every charge displaced N(0, 0.02 A), and
    Poisson arrivals every mean_gap_ms on average.  Returns (sizes, nsteps,
    arrivals_s, charges, spec) with charges[k] = [nsteps[k], 30, 4] and
    spec(k) building request k's system (seconds of host time for the largest)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(500), np.log(40000), size=nreq)).astype(int)
    nsteps = rng.integers(10, 21, size=nreq)
    gaps = rng.exponential(mean_gap_ms * 1e-3, size=nreq)
    arrivals = np.cumsum(gaps) - gaps[0]
    charges = []
    for k in range(nreq):
        base = qm_charge_sets(1, 2000 + k)[0]
        steps = np.repeat(base[None], nsteps[k], axis=0)
        steps[:, :, :3] += rng.normal(0.0, 0.02 * ANG2AU, size=(nsteps[k], base.shape[0], 3))
        charges.append(steps)

    def spec(k):
        radius = (3.0 * sizes[k] / (4.0 * np.pi * 0.1)) ** (1.0 / 3.0) + 3.0
        return sphere_system(int(sizes[k]), radius, seed=1000 + k, cavity=6.0)
    return sizes, nsteps, arrivals, charges, spec
