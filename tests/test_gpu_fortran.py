"""The Fortran ISO_C_BINDING module drives the GPU solve: a Fortran program
(tests/fortran/test_fortran_api.F90, built by __graft_entry__.build) calls
ommp_init_mmp / ommp_set_external_field / ommp_get_polelec_energy, reads the
dipoles through the module, evaluates external-point potentials and fields and
the polarization-energy gradient; results must equal the oracle's."""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

import mmp_io
import oracle_api as oa

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "openmmpol_amd", "fortran", "test_fortran_api")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (flang missing at build time)")
@pytest.mark.parametrize("name,ff,k", [("N-methylacetamide", "AMOEBA", 1), ("1crn", "WANG_AL", 1),
                                       ("1crn", "AMOEBA", 0)])
def test_fortran_module_solve(name, ff, k, fixture_file, tmp_path):
    mmp = fixture_file(name, f"input_{ff}.mmp")
    eff = "none"
    if k:
        eff = str(tmp_path / "ef.txt")
        with gzip.open(mmp_io.fixture_path(name, "EF_1.txt"), "rb") as fi, open(eff, "wb") as fo:
            shutil.copyfileobj(fi, fo)
    r = subprocess.run([EXE, mmp, eff], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    ep = float(next(l for l in lines if l.startswith("EP")).split()[1])
    ipd = np.array([[float(v) for v in l.split()[1:]] for l in lines if l.startswith("IPD")])
    o = oa.OracleSystem(mmp_io.prepare_mmp(mmp_io.fixture_path(name, f"input_{ff}.mmp")))
    ext = np.zeros((o.npol, 3)) if k == 0 else mmp_io.read_ef(mmp_io.fixture_path(name, "EF_1.txt"))[o.polar_mm]
    o.set_external_field(ext)
    eo = o.get_polelec_energy()
    assert abs(ep - eo) <= 1e-10 * abs(eo)
    em = float(next(l for l in lines if l.startswith("EM")).split()[1])
    emo = o.energy_mm()
    assert abs(em - emo) <= 1e-12 * abs(emo)
    assert np.abs(ipd - o.ipd.reshape(-1, 3)).max() < 1e-9
    # v = V_mmpol + V_mm, e = E_mmpol + E_pol at two fixed points
    x2e = np.array([[float(v) for v in l.split()[1:]] for l in lines if l.startswith("X2E")])
    cext = np.array([[25.0, -31.0, 17.0], [-40.0, 12.5, -22.0]])
    vmp, emp = o.x2e(cext, 3)
    vm, _ = o.x2e(cext, 1)
    _, ep_ = o.x2e(cext, 2)
    assert np.allclose(x2e[:, 0], vmp + vm, rtol=1e-12, atol=1e-15)
    assert np.allclose(x2e[:, 1:], emp + ep_, rtol=1e-10, atol=1e-14)
    # gradient after the solve; the oracle's own dipoles differ by < 1e-9 a.u.
    grd = np.array([[float(v) for v in l.split()[1:]] for l in lines if l.startswith("GRD")])
    go = o.polelec_geomgrad()
    assert grd.shape == go.shape
    assert np.abs(grd - go).max() <= 1e-6 * np.abs(go).max()
