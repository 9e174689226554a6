"""Potentials and fields at external points (ommp_potential_* / ommp_field_*,
mod_interface.F90:250-341) on the GPU against the oracle's restatement of
potential_M2E / potential_D2E / field_M2E / field_D2E: 1e-12 relative to the
largest value (FP64 sums of the same terms in another order)."""
import ctypes

import numpy as np
import pytest

import mmp_io
import oracle_api as oa

pytestmark = pytest.mark.gpu


def _points(cmm, n, rng, margin=1.5):
    c = np.asarray(cmm).reshape(-1, 3)
    lo, hi = c.min(0) - 4.0, c.max(0) + 4.0
    pts = []
    while len(pts) < n:
        p = rng.uniform(lo, hi, size=(4 * n, 3))
        for q in p:
            if len(pts) < n and np.min(np.linalg.norm(c - q, axis=1)) > margin:
                pts.append(q)
    return np.array(pts)


@pytest.mark.parametrize("name,ff", [("N-methylacetamide", "AMOEBA"), ("N-methylacetamide", "WANG_AL"),
                                     ("1crn", "AMOEBA"), ("1crn", "WANG_AL")])
def test_external_potentials_and_fields(name, ff, fixture_file):
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file(name, f"input_{ff}.mmp"))
    o = oa.OracleSystem(mmp_io.prepare_mmp(mmp_io.fixture_path(name, f"input_{ff}.mmp")))
    ext = mmp_io.read_ef(mmp_io.fixture_path(name, "EF_1.txt"))[np.asarray(s.polar_mm)]
    s.set_external_field(ext)
    o.set_external_field(ext)
    pts = _points(np.asarray(s.cmm), 300, np.random.default_rng(7))
    for which, pot, fld in ((1, s.potential_mm2ext, s.field_mm2ext), (2, s.potential_pol2ext, s.field_pol2ext),
                            (3, s.potential_mmpol2ext, s.field_mmpol2ext)):
        # the GPU dipoles, so that the comparison isolates the external-point kernels
        V, E = o.x2e(pts, which, s.ipd.copy())
        assert np.abs(pot(pts) - V).max() <= 1e-12 * np.abs(V).max(), which
        assert np.abs(fld(pts) - E).max() <= 1e-12 * np.abs(E).max(), which
    s.terminate()


def test_results_are_added_to_the_callers_arrays(fixture_file):
    from openmmpol_amd import OMMPSystem
    from openmmpol_amd._lib import lib
    s = OMMPSystem(fixture_file("1crn", "input_AMOEBA.mmp"))
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    pts = _points(np.asarray(s.cmm), 50, np.random.default_rng(8))
    dp = ctypes.POINTER(ctypes.c_double)
    v0 = np.random.default_rng(1).normal(size=50)
    e0 = np.random.default_rng(2).normal(size=(50, 3))
    v, e = v0.copy(), e0.copy()
    lib().ommp_potential_mmpol2ext(s.handler, 50, pts.ctypes.data_as(dp), v.ctypes.data_as(dp))
    lib().ommp_field_mmpol2ext(s.handler, 50, pts.ctypes.data_as(dp), e.ctypes.data_as(dp))
    assert np.allclose(v, v0 + s.potential_mmpol2ext(pts), rtol=1e-14, atol=1e-14)
    assert np.allclose(e, e0 + s.field_mmpol2ext(pts), rtol=1e-14, atol=1e-14)
    s.terminate()


def test_external_potential_on_a_large_grid():
    """10,000 grid-like points around the 15,000-site sphere (a QM-grid-sized
    request), multipoles and dipoles."""
    from openmmpol_amd import synth
    spec = synth.sphere_system(15000, 33.0, seed=2, cavity=6.0)
    s = synth.make_system(spec)
    s.set_precision("mixed")
    s.set_solver_tolerance(3e-10, 500)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    rng = np.random.default_rng(3)
    d = rng.normal(size=(10000, 3))
    pts = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0.0, 5.0 * 1.8897261245650, size=(10000, 1))
    o = oa.OracleSystem(synth.oracle_prep(spec))
    V, E = o.x2e(pts, 3, s.ipd.copy())
    assert np.abs(s.potential_mmpol2ext(pts) - V).max() <= 1e-12 * np.abs(V).max()
    assert np.abs(s.field_mmpol2ext(pts) - E).max() <= 1e-12 * np.abs(E).max()
    s.terminate()
