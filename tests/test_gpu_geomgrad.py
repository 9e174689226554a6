"""Geometric gradient of the polarization energy on the GPU (ommp_polelec_geomgrad,
mod_geomgrad.F90:113-273; ommp_rotation_geomgrad, rotate_multipoles.F90:3-230).

- against the oracle's restatement, fed the GPU's own dipoles so that the
  comparison isolates the gradient kernels: 1e-11 of the largest component
  (FP64 sums of the same terms in another order), reference fixtures (AMOEBA
  and Wang-AL), a cluster with every frame convention, a 3,000-site sphere,
  7,001-site spheres (AMOEBA, Wang-AL) where the dipole-dipole part runs
  pair-once;
- against the 'Grad EP' listings of the reference's FULL_POTENTIAL.ref
  (kcal/mol/angstrom), as the oracle is;
- at the 15,000-site benchmark size, through properties: the gradient sums to
  zero (traceless quadrupoles: the energy is translation invariant) and its
  projection on a random displacement equals the central difference of the
  energy along it;
- the API semantics: solve-on-demand, numerical=True, the rotation entry point.
"""
import numpy as np
import pytest

import mmp_io
import oracle_api as oa

pytestmark = pytest.mark.gpu

CONV = mmp_io.AU2KCALMOL * mmp_io.ANG2AU


def _oracle_grad(prep, ipd):
    return oa.OracleSystem(prep).polelec_geomgrad(ipd)


@pytest.mark.parametrize("name,ff", [("N-methylacetamide", "AMOEBA"), ("N-methylacetamide", "WANG_AL"),
                                     ("1crn", "AMOEBA"), ("1crn", "WANG_AL"), ("1ubq", "AMOEBA")])
def test_gradient_matches_oracle_on_fixtures(name, ff, fixture_file):
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file(name, f"input_{ff}.mmp"))
    ext = mmp_io.read_ef(mmp_io.fixture_path(name, "EF_1.txt"))[np.asarray(s.polar_mm)]
    s.set_external_field(ext)
    g = s.polelec_geomgrad()
    want = _oracle_grad(mmp_io.prepare_mmp(mmp_io.fixture_path(name, f"input_{ff}.mmp")), s.ipd.copy())
    assert g.shape == (s.mm_atoms, 3)
    assert np.abs(g - want).max() <= 1e-11 * np.abs(want).max()
    s.terminate()


@pytest.mark.parametrize("name,tol", [("N-methylacetamide", 2e-6), ("1crn", 5e-6), ("1ubq", 2e-6)])
def test_gradient_matches_reference_listing(name, tol, fixture_file):
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file(name, "input_AMOEBA.mmp"))
    s.set_solver_tolerance(1e-12, 500)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    full = mmp_io.fixture_path(name, "FULL_POTENTIAL.ref")
    want = mmp_io.read_grad(full, "EP")
    got = s.polelec_geomgrad() * CONV
    assert np.abs(got - want).max() <= tol * np.abs(want).max()
    s.terminate()


def test_every_frame_convention_matches_oracle():
    from openmmpol_amd import OMMPSystem
    from test_oracle_geomgrad import _frames_cluster
    prep = _frames_cluster()
    q = prep["q0"].copy()
    q[:, 4:10] *= 3.0  # the .mmp convention
    s = OMMPSystem.from_arrays(True, prep["cmm"], q, prep["pol"], prep["adj_ri"], prep["adj_ci"],
                               pgroup=prep["grp"], frames=prep["frames"])
    s.set_solver_tolerance(1e-12, 500)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    g = s.polelec_geomgrad()
    want = _oracle_grad(prep, s.ipd.copy())
    assert np.abs(g - want).max() <= 1e-11 * np.abs(want).max()
    s.terminate()


def test_rotation_entry_point(fixture_file):
    from openmmpol_amd import OMMPSystem
    prep = mmp_io.prepare_mmp(mmp_io.fixture_path("1crn", "input_AMOEBA.mmp"))
    s = OMMPSystem(fixture_file("1crn", "input_AMOEBA.mmp"))
    rng = np.random.default_rng(5)
    E = rng.normal(0.0, 0.01, (s.mm_atoms, 3))
    G = rng.normal(0.0, 0.01, (s.mm_atoms, 6))
    got = s.rotation_geomgrad(E, G)
    want = oa.rotation_geomgrad(prep["cmm"], prep["q0"], prep["frames"], E, G)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    s.terminate()
    w = OMMPSystem(fixture_file("1crn", "input_WANG_AL.mmp"))
    assert not w.rotation_geomgrad(E, G).any()
    w.terminate()


def test_gradient_solves_first_and_numerical_option(fixture_file):
    from openmmpol_amd import OMMPSystem
    path = fixture_file("N-methylacetamide", "input_AMOEBA.mmp")
    a = OMMPSystem(path)
    a.set_solver_tolerance(1e-11, 500)
    g_fresh = a.polelec_geomgrad()  # no solve yet: default solver, no external field
    b = OMMPSystem(path)
    b.set_solver_tolerance(1e-11, 500)
    b.set_external_field(np.zeros((b.pol_atoms, 3)))
    g = b.polelec_geomgrad()
    assert np.abs(g_fresh - g).max() <= 1e-13 * np.abs(g).max()
    # Central differences (step 1e-5 bohr) of energies from the reference's CG,
    # whose small-step exit (|p.Tp| < eps_rp) ends these solves near 1e-9
    # relative whatever the tolerance: the oracle's differences of such solves
    # agree with the analytic gradient to 1.0e-5 relative, and to 1.5e-10 with
    # the exit switched off.  The reference's own test_SI_geomgrad vs _num
    # comparison uses rtol 1e-3.
    c0 = np.array(b.cmm)
    num = b.polelec_geomgrad(numerical=True)
    assert np.array_equal(np.array(b.cmm), c0)
    assert np.abs(num - g).max() <= 3e-5 * np.abs(g).max()
    a.terminate()
    b.terminate()


def test_sphere3000_matches_oracle():
    from openmmpol_amd import synth
    spec = synth.sphere_system(3000, 19.0, seed=4)
    s = synth.make_system(spec)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    g = s.polelec_geomgrad()
    want = _oracle_grad(synth.oracle_prep(spec), s.ipd.copy())
    assert np.abs(g - want).max() <= 1e-11 * np.abs(want).max()
    s.terminate()


@pytest.mark.parametrize("ff", ["AMOEBA", "WANG_AL"])
def test_sphere7001_pair_once_matches_oracle(ff):
    """From 5,000 polarizable sites the dipole-dipole part is evaluated once per
    unordered pair (its term added at one site, subtracted at the other); 7,001
    sites (27 x 256 + 89) leave a partial last block."""
    from openmmpol_amd import synth
    n = 7001
    spec = dict(synth.sphere_system(n, (3.0 * n / (4.0 * np.pi * 0.0996)) ** (1.0 / 3.0), seed=24))
    if ff == "WANG_AL":
        spec.update(amoeba=False, q=spec["q"][:, :1].copy(), pgroup=None, scales=None)
    s = synth.make_system(spec)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    g = s.polelec_geomgrad()
    prep = synth.oracle_prep(spec)
    if ff == "WANG_AL":
        prep["scales"] = mmp_io.SCALES_WANG_AL.copy()
        prep["thole_scale"] = mmp_io.THOLE_SCALE_WANG_AL
    want = _oracle_grad(prep, s.ipd.copy())
    err = np.abs(g - want).max() / np.abs(want).max()
    print(f"{ff} {n} sites: gradient err {err:.2e} of the largest component", flush=True)
    assert err <= 1e-11
    s.terminate()


def test_sphere9000_nonpolarizable_sources_matches_oracle():
    """From 5,000 polarizable sites the static part evaluates each (polarizable,
    MM) pair once for both ends, the MM end taking minus the polarizable end's
    term.  Every seventh site non-polarizable (9,000 MM sites, 7,714
    polarizable): MM-only sources and targets, two block counts, partial last
    blocks on both ends."""
    from openmmpol_amd import synth
    n = 9000
    spec = dict(synth.sphere_system(n, (3.0 * n / (4.0 * np.pi * 0.0996)) ** (1.0 / 3.0), seed=25))
    spec["pol"] = spec["pol"].copy()
    spec["pol"][::7] = 0.0
    s = synth.make_system(spec)
    assert s.pol_atoms == n - len(range(0, n, 7)) and s.pol_atoms >= 5000
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    g = s.polelec_geomgrad()
    want = _oracle_grad(synth.oracle_prep(spec), s.ipd.copy())
    err = np.abs(g - want).max() / np.abs(want).max()
    print(f"AMOEBA {n} MM sites, {s.pol_atoms} polarizable: gradient err {err:.2e} of the largest component",
          flush=True)
    assert err <= 1e-11
    s.terminate()


def test_sphere15000_translation_and_directional_derivative():
    from openmmpol_amd import synth
    spec = synth.sphere_system(15000, 33.0, seed=2)
    s = synth.make_system(spec)
    s.set_precision("fp64")
    s.set_solver_tolerance(1e-11, 500)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    g = s.polelec_geomgrad()
    gmax = np.abs(g).max()
    assert gmax > 0
    assert np.abs(g.sum(axis=0)).max() <= 1e-12 * gmax * len(g)
    c0 = np.array(s.cmm)
    v = np.random.default_rng(11).normal(size=c0.shape)
    v /= np.linalg.norm(v)
    h = 1e-3
    e = []
    for sign in (1.0, -1.0):
        s.update_coordinates(c0 + sign * h * v)
        s.set_external_field(np.zeros((s.pol_atoms, 3)))
        e.append(s.get_polelec_energy())
    fd = (e[0] - e[1]) / (2 * h)
    an = float(np.sum(g * v))
    assert abs(fd - an) <= 1e-6 * np.linalg.norm(g)
    s.terminate()
