"""GPU parity: the HIP library through its C-ABI against the FP64 oracle on the
same inputs, and against the reference's own fixtures.

Tolerances: FP64 path vs oracle (same algorithm, same iterate sequence up to
rounding): dipoles 1e-9 a.u. max-abs, energies 1e-10 relative.  Mixed path vs
an exact dense solve: dipoles 1e-7 a.u., energy 1e-9 relative.  Reference
fixtures: the reference's own ctest tolerances (IPD rtol 1e-6 atol 1e-5,
energy 1e-6).
"""
import numpy as np
import pytest

import mmp_io
import oracle_api as oa

pytestmark = pytest.mark.gpu

FIXTURES = [("N-methylacetamide", "AMOEBA"), ("N-methylacetamide", "WANG_AL"),
            ("N-methylacetamide", "AMOEBA_CUT"), ("N-methylacetamide", "WANG_AL_CUT"),
            ("1crn", "AMOEBA"), ("1crn", "WANG_AL")]


def _ext(name, polar_mm, k):
    if k == 0:
        return np.zeros((len(polar_mm), 3))
    return mmp_io.read_ef(mmp_io.fixture_path(name, "EF_1.txt"))[np.asarray(polar_mm)]


def _pair(name, ff, fixture_file):
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file(name, f"input_{ff}.mmp"))
    o = oa.OracleSystem(mmp_io.prepare_mmp(mmp_io.fixture_path(name, f"input_{ff}.mmp")))
    return s, o


@pytest.mark.parametrize("name,ff", FIXTURES)
@pytest.mark.parametrize("k", [0, 1])
def test_set_external_field_fp64(name, ff, k, fixture_file):
    s, o = _pair(name, ff, fixture_file)
    ext = _ext(name, s.polar_mm, k)
    s.set_external_field(ext)
    ipd = s.ipd.copy()
    ep = s.get_polelec_energy()
    o.set_external_field(ext)
    eo = o.get_polelec_energy()
    assert np.abs(ipd - o.ipd).max() < 1e-9
    assert abs(ep - eo) <= 1e-10 * abs(eo)
    it = s.last_iterations
    assert all(abs(it[m] - o.iters[m]) <= 1 for m in range(o.n_ipd))
    ref = mmp_io.read_ipd(mmp_io.fixture_path(name, f"IPD_{k}_{ff}.ref"))
    assert np.allclose(ipd.reshape(-1, 3), ref, rtol=1e-6, atol=1e-5)
    if "CUT" not in ff:
        eref = mmp_io.read_ene(mmp_io.fixture_path(name, f"ENE_{k}_{ff}.ref"))["EP"]
        # np.isclose(rtol=1e-6, atol=1e-6) as compare_potential.py:53 with the ctest arguments
        assert abs(ep - eref) <= 1e-6 + 1e-6 * abs(eref)


@pytest.mark.parametrize("name,ff", [("1crn", "AMOEBA"), ("1crn", "WANG_AL"), ("1ubq", "AMOEBA")])
def test_static_field_matches_oracle(name, ff, fixture_file):
    s, o = _pair(name, ff, fixture_file)
    e = s.e_m2d
    eo = o.m2d()
    assert np.abs(e - eo).max() <= 1e-12 * np.abs(eo).max()


@pytest.mark.parametrize("name,ff", [("1crn", "AMOEBA"), ("1crn", "WANG_AL"), ("N-methylacetamide", "WANG_AL")])
def test_tmatvec_kinds(name, ff, fixture_file):
    s, o = _pair(name, ff, fixture_file)
    rng = np.random.default_rng(4)
    x = rng.normal(0.0, 0.05, size=(2, s.pol_atoms, 3))
    ref = np.stack([o.tmatvec(x[0]), o.tmatvec(x[1])])
    dense = (o.tmat() @ x.reshape(2, -1).T).T.reshape(2, -1, 3)
    scale = np.abs(ref).max()
    for kind, tol in (("otf64", 1e-13), ("incore64", 1e-13), ("otf32", 2e-6), ("incore32", 2e-6),
                      ("sym64", 1e-13), ("plain64", 1e-13), ("sym32", 2e-6), ("plain32", 2e-6)):
        for ncols in (1, 2):
            y, _ = s.tmatvec(x[:ncols], kind=kind)
            want = dense if kind.startswith("incore") else ref
            assert np.abs(y - want[:ncols]).max() <= tol * scale, (kind, ncols)


@pytest.mark.parametrize("ff", ["AMOEBA", "WANG_AL"])
@pytest.mark.parametrize("solver,matv", [("cg", "incore"), ("diis", "direct"), ("diis", "incore"),
                                         ("inversion", "direct")])
def test_solver_and_matv_variants(ff, solver, matv, fixture_file):
    s, o = _pair("1crn", ff, fixture_file)
    s.set_default_matv(matv)
    s.set_external_field(np.zeros((s.pol_atoms, 3)), solver=solver)
    ipd = s.ipd.copy()
    E = o.m2d()
    sid = {"cg": oa.SOLVER_CG, "diis": oa.SOLVER_DIIS, "inversion": oa.SOLVER_INVERSION}[solver]
    mid = {"direct": oa.MATV_DIRECT, "incore": oa.MATV_INCORE}[matv]
    want = o.polarization(E, solver=sid, matv=mid, use_guess=False)
    tol = 1e-11 if solver == "inversion" else 1e-9
    assert np.abs(ipd - want).max() < tol


@pytest.mark.parametrize("ff", ["AMOEBA", "WANG_AL"])
@pytest.mark.parametrize("matv", ["direct", "incore"])
def test_mixed_precision_against_exact_solve(ff, matv, fixture_file):
    s, o = _pair("1crn", ff, fixture_file)
    ext = _ext("1crn", s.polar_mm, 1)
    s.set_precision("mixed")
    s.set_default_matv(matv)
    s.set_solver_tolerance(1e-10)
    s.set_external_field(ext)
    ipd = s.ipd.copy()
    ep = s.get_polelec_energy()
    E = o.m2d() + ext[None]
    T = o.tmat()
    exact = np.stack([np.linalg.solve(T, E[m].ravel()).reshape(-1, 3) for m in range(o.n_ipd)])
    ee = o.energy(exact)
    assert np.abs(ipd - exact).max() < 1e-7
    assert abs(ep - ee) <= 1e-9 * abs(ee)
    outer, inner = s.last_iterations[2], s.last_iterations[3]
    assert 1 <= outer <= 6 and inner > 0


def test_warm_start_and_coordinate_update(fixture_file):
    """ipd_use_guess after a solve; update_coordinates clears it and the static
    field (mod_polarization.F90:192-210, 305-307; mod_mmpol.F90:325-372)."""
    from openmmpol_amd import OMMPSystem
    s, o = _pair("1crn", "AMOEBA", fixture_file)
    ext = _ext("1crn", s.polar_mm, 1)
    s.set_external_field(ext)
    o.set_external_field(ext)
    it_cold = s.last_iterations[:2]
    s.set_external_field(ext)  # warm start from the converged dipoles
    o.set_external_field(ext)
    assert max(s.last_iterations[:2]) < max(it_cold)
    assert np.abs(s.ipd - o.ipd).max() < 1e-9
    rng = np.random.default_rng(2)
    newc = s.cmm.copy() + rng.normal(0.0, 0.02, size=s.cmm.shape)
    s.update_coordinates(newc)
    s.set_external_field(ext)
    prep = mmp_io.prepare_mmp(mmp_io.fixture_path("1crn", "input_AMOEBA.mmp"))
    prep["cmm"] = newc
    prep["q"] = mmp_io.rotate_multipoles(newc, prep["q0"], prep["frames"])
    o2 = oa.OracleSystem(prep)
    o2.set_external_field(ext)
    assert np.abs(s.ipd - o2.ipd).max() < 1e-9
    assert abs(s.get_polelec_energy() - o2.get_polelec_energy()) <= 1e-10 * abs(o2.get_polelec_energy())


@pytest.mark.parametrize("matv", ["direct", "incore"])
def test_nomm_solves_d_set_only(matv, fixture_file):
    s, o = _pair("1crn", "AMOEBA", fixture_file)
    ext = _ext("1crn", s.polar_mm, 1)
    s.set_external_field(ext, nomm=True, matv=matv)
    o.set_external_field(ext, nomm=True, matv={"direct": oa.MATV_DIRECT, "incore": oa.MATV_INCORE}[matv])
    ipd = s.ipd.copy()
    assert np.abs(ipd[0] - o.ipd[0]).max() < 1e-9
    assert np.abs(ipd[1]).max() == 0.0


def test_energy_getter_solves_when_needed(fixture_file):
    s, o = _pair("1crn", "WANG_AL", fixture_file)
    ep = s.get_polelec_energy()
    eo = o.get_polelec_energy()
    assert abs(ep - eo) <= 1e-10 * abs(eo)


@pytest.mark.parametrize("n,radius", [(3000, 19.0), (15000, 33.0)])
def test_pair_once_matvec(n, radius):
    """The pair-once sweep (each unordered pair evaluated once, both directions
    accumulated) against the oracle's matvec at 3,000 sites (12 blocks of 256,
    the last one padded) and against the per-target sweep at 15,000; both
    precisions, one and two columns, bitwise repeatable."""
    from openmmpol_amd import synth
    spec = synth.sphere_system(n, radius, seed=6)
    s = synth.make_system(spec)
    x = np.random.default_rng(8).normal(0.0, 0.05, size=(2, s.pol_atoms, 3))
    if n <= 3000:
        o = oa.OracleSystem(synth.oracle_prep(spec))
        ref = np.stack([o.tmatvec(x[0]), o.tmatvec(x[1])])
    else:
        ref, _ = s.tmatvec(x, kind="plain64")
    scale = np.abs(ref).max()
    for kind, tol in (("sym64", 1e-13), ("sym32", 2e-6)):
        for ncols in (1, 2):
            y, _ = s.tmatvec(x[:ncols], kind=kind)
            assert np.abs(y - ref[:ncols]).max() <= tol * scale, (kind, ncols)
        y2, _ = s.tmatvec(x, kind=kind)
        assert np.array_equal(y2, s.tmatvec(x, kind=kind)[0])
    s.terminate()


def test_pair_once_matvec_mixed_thole():
    """AMOEBA with a Thole a that differs between sites (damping with
    min(a_i, a_j)) at 13,000 sites: the pair-once sweep in both precisions
    against the per-target FP64 sweep, and a visible change from the uniform
    a = 0.39 of the other synthetic systems."""
    from openmmpol_amd import synth
    spec = synth.sphere_system(13000, 31.5, seed=14)
    x = np.random.default_rng(15).normal(0.0, 0.05, size=(2, 13000, 3))
    ys = {}
    for name, thole in (("uniform", None), ("mixed", np.random.default_rng(16).uniform(0.3, 0.45, 13000))):
        sp = dict(spec)
        if thole is not None:
            sp["thole_a"] = thole
        s = synth.make_system(sp)
        want, _ = s.tmatvec(x, kind="plain64")
        for kind, tol in (("sym64", 1e-13), ("sym32", 2e-6)):
            for ncols in (1, 2):
                got, _ = s.tmatvec(x[:ncols], kind=kind)
                assert np.abs(got - want[:ncols]).max() <= tol * np.abs(want).max(), (name, kind, ncols)
        ys[name] = want
        s.terminate()
    assert np.abs(ys["mixed"] - ys["uniform"]).max() > 1e-6 * np.abs(ys["uniform"]).max()


def test_pair_once_matvec_amber():
    """The AMBER (Wang linear Thole, one dipole set) form of the pair-once
    sweep against the per-target sweep at 13,000 sites (both precisions; the
    per-target kernel is pinned to the oracle on the Wang-AL fixtures)."""
    from openmmpol_amd import OMMPSystem, synth
    spec = synth.sphere_system(13000, 31.5, seed=12)
    s = OMMPSystem.from_arrays(False, spec["cmm"], spec["q"][:, :1], spec["pol"], spec["bonds_ri"],
                               spec["bonds_ci"])
    x = np.random.default_rng(13).normal(0.0, 0.05, size=(1, s.pol_atoms, 3))
    want, _ = s.tmatvec(x, kind="plain64")
    for kind, tol in (("sym64", 1e-13), ("sym32", 2e-6), ("plain32", 2e-6)):
        got, _ = s.tmatvec(x, kind=kind)
        assert np.abs(got - want).max() <= tol * np.abs(want).max(), kind
    s.terminate()


@pytest.mark.parametrize("ff", ["AMOEBA", "WANG_AL"])
def test_turn_pol_off_solves_match_oracle(ff, fixture_file):
    """ommp_turn_pol_off after a solve (device state built for the old set):
    FP64 and mixed solves of the reduced set against the oracle's, the
    fixed-multipole energy unchanged (it does not involve polarizabilities)."""
    s, o = _pair("1crn", ff, fixture_file)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    e_mm = s.get_fixedelec_energy()
    pm = o.polar_mm
    off = [int(pm[k]) + 1 for k in (0, 5, 6, 100, 101, 102, 300)] + [int(pm[-1]) + 1]
    s.turn_pol_off(off)
    o.turn_pol_off(off)
    assert s.pol_atoms == o.npol and np.array_equal(s.polar_mm, o.polar_mm)
    ef = mmp_io.read_ef(mmp_io.fixture_path("1crn", "EF_1.txt"))[o.polar_mm]
    # FP64 against the oracle's CG at the same tolerance (same iterates)
    o.ipd_use_guess = False
    o.set_external_field(ef)
    eo = o.get_polelec_energy()
    s.set_external_field(ef)
    e = s.get_polelec_energy()
    print(f"{ff} fp64: {s.pol_atoms} sites left, E_pol {e:.12e} oracle {eo:.12e}", flush=True)
    assert np.abs(s.ipd - o.ipd).max() < 1e-9
    assert abs(e - eo) <= 1e-10 * abs(eo)
    # mixed against the exact dense solve of the reduced set
    E = o.m2d() + ef[None]
    T = o.tmat()
    exact = np.stack([np.linalg.solve(T, E[m].ravel()).reshape(-1, 3) for m in range(o.n_ipd)])
    ee = o.energy(exact)
    s.set_precision("mixed")
    s.set_solver_tolerance(1e-10, 500)
    s.reset_guess(False)
    s.set_external_field(ef)
    e = s.get_polelec_energy()
    print(f"{ff} mixed: E_pol {e:.12e} exact {ee:.12e}", flush=True)
    assert np.abs(s.ipd - exact).max() < 1e-7
    assert abs(e - ee) <= 1e-9 * abs(ee)
    assert s.get_fixedelec_energy() == e_mm
    s.terminate()


def test_turn_pol_off_every_site(fixture_file):
    """The reference's own use (tests/test_programs/C/test_SI_potential.c:88-95):
    every atom of a system switched off; no polarizable site is left and the
    polarization energy is zero."""
    s, o = _pair("N-methylacetamide", "AMOEBA", fixture_file)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    s.turn_pol_off(np.arange(1, s.mm_atoms + 1))
    assert s.pol_atoms == 0
    s.set_external_field(np.zeros((0, 3)))
    assert s.get_polelec_energy() == 0.0
    s.terminate()


def test_vector_fma_rate_is_plausible():
    """ommp_amd_vector_fma_tflops: the FMA rate of non-packed vector code, the
    reference point bench.py reports beside the pair kernels' rooflines; between
    1 TFLOP/s and twice the published vector peaks (FP64 78.6, FP32 157.3)."""
    from openmmpol_amd._lib import lib
    r64 = lib().ommp_amd_vector_fma_tflops(64)
    r32 = lib().ommp_amd_vector_fma_tflops(32)
    print(f"non-packed FMA rate: FP32 {r32:.1f}, FP64 {r64:.1f} TFLOP/s", flush=True)
    assert 1.0 < r64 < 2 * 78.6
    assert 1.0 < r32 < 2 * 157.3
