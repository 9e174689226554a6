"""CPU-side checks of the C-ABI library: it loads, exports every declared
symbol, and its host-side preparation (.mmp reader, rotation, topology,
screening lists) equals the oracle's and the reference's summary fixtures.
No device work happens in this file."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import mmp_io
import oracle_api as oa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = []
    for h in ("openmmpol.h", "ommp_amd.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"\b(ommp_\w+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    import openmmpol_amd
    L = openmmpol_amd.lib()
    missing = [n for n in _declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    assert len(_declared_symbols()) >= 40


def test_bindings_cover_declared_symbols():
    from openmmpol_amd._lib import SIGNATURES
    bound = {s[0] for s in SIGNATURES}
    assert set(_declared_symbols()) <= bound


CASES = [("N-methylacetamide", "AMOEBA"), ("N-methylacetamide", "WANG_AL"), ("N-methylacetamide", "AMOEBA_CUT"),
         ("1crn", "AMOEBA"), ("1crn", "WANG_AL"), ("1ubq", "AMOEBA")]


def _csr_dict(ri, ci, val, nrow):
    return {(i, int(ci[k])): float(val[k]) for i in range(nrow) for k in range(ri[i], ri[i + 1])}


@pytest.mark.parametrize("name,ff", CASES)
def test_host_preparation_matches_oracle(name, ff, fixture_file):
    from openmmpol_amd import OMMPSystem
    path = fixture_file(name, f"input_{ff}.mmp")
    s = OMMPSystem(path)
    prep = mmp_io.prepare_mmp(mmp_io.fixture_path(name, f"input_{ff}.mmp"))
    o = oa.OracleSystem(prep)
    assert s.pol_atoms == o.npol and s.mm_atoms == prep["nmm"]
    assert s.n_ipd == (2 if prep["amoeba"] else 1)
    assert np.array_equal(s.polar_mm, o.polar_mm)
    assert np.array_equal(s.cmm, prep["cmm"])
    assert np.abs(s._q - prep["q"]).max() < 1e-14
    assert np.abs(s.thole - o.thole()).max() < 1e-15
    assert np.array_equal(s.cpol, prep["cmm"][o.polar_mm])
    for which in range(0, 7):
        if which == 1 and not prep["amoeba"]:
            continue
        nrow = s.pol_atoms if which == 2 else s.mm_atoms
        a = _csr_dict(*s.get_list(which), nrow)
        b = _csr_dict(*o.list_csr(which), nrow)
        assert a == b, which
    s.terminate()


def test_atomic_numbers_are_zero_from_mmp(fixture_file):
    """ommp_get_zmm (mod_c_interface.F90:950-959): the .mmp loader skips the
    atomic-number section (mod_inputloader.F90:131) of an array that
    topology_init zeroed, so the getter returns mm_atoms zeros."""
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file("1crn", "input_AMOEBA.mmp"))
    assert s.zmm.shape == (s.mm_atoms,) and not s.zmm.any()
    s.terminate()


def test_host_preparation_matches_summary(fixture_file):
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file("1crn", "input_AMOEBA.mmp"))
    summ = mmp_io.read_summary(mmp_io.fixture_path("1crn", "summary_AMOEBA.ref"))
    assert np.abs(s.cmm - summ["cmm"]).max() < 1e-8
    assert np.abs(s._q - summ["q"]).max() < 1e-8
    assert np.abs(s.pol - summ["pol"][:, 0]).max() < 1e-8
    assert np.abs(s.thole - summ["thole"][:, 0]).max() < 1e-8


def _run(code):
    return subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)


def test_fatal_errors_exit_1_like_the_reference():
    """fatal_error prints '[stop]' and exit(1) (mod_io.F90:155-175)."""
    r = _run("import openmmpol_amd as o; o.OMMPSystem('/nonexistent.mmp')")
    assert r.returncode == 1 and "[stop]" in r.stdout and "does not exist" in r.stdout
    r = _run("import openmmpol_amd as o, numpy as np\n"
             "s = o.OMMPSystem.from_arrays(True, np.zeros((1,3)), np.zeros((1,10)), [1.0], [0,0], [], pgroup=[0])\n"
             "o.lib().ommp_set_default_solver(s.handler, 7)")
    assert r.returncode == 1 and "Unrecognized setting for default solver method" in r.stdout
    r = _run("import openmmpol_amd as o, numpy as np\n"
             "s = o.OMMPSystem.from_arrays(True, np.zeros((1,3)), np.zeros((1,10)), [1.0], [0,0], [], pgroup=[0])\n"
             "o.lib().ommp_set_default_matv(s.handler, 9)")
    assert r.returncode == 1 and "Unrecognized setting for default matrix-vector method" in r.stdout
    # potential_D2E / field_D2E need a solve first (mod_electrostatics.F90:3090, 3238)
    r = _run("import openmmpol_amd as o, numpy as np\n"
             "s = o.OMMPSystem.from_arrays(True, np.zeros((1,3)), np.zeros((1,10)), [1.0], [0,0], [], pgroup=[0])\n"
             "s.potential_pol2ext(np.ones((2, 3)))")
    assert r.returncode == 1 and "IPD should be computed before D2E potential" in r.stdout
    r = _run("import openmmpol_amd as o, numpy as np\n"
             "s = o.OMMPSystem.from_arrays(True, np.zeros((1,3)), np.zeros((1,10)), [1.0], [0,0], [], pgroup=[0])\n"
             "s.field_mmpol2ext(np.ones((2, 3)))")
    assert r.returncode == 1 and "IPD should be computed before computing D2E field" in r.stdout


def test_python_argument_checks(fixture_file):
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file("N-methylacetamide", "input_AMOEBA.mmp"))
    with pytest.raises(ValueError):
        s.set_external_field(np.zeros((3, 3)))
    with pytest.raises(ValueError):
        s.set_external_field(np.zeros((s.pol_atoms, 3)), solver="lu")
    with pytest.raises(ValueError):
        s.set_external_field(np.zeros((s.pol_atoms, 3)), matv="dense")
    with pytest.raises(ValueError):
        s.update_coordinates(np.zeros((2, 3)))


def test_synthetic_generators_are_deterministic():
    from openmmpol_amd import synth
    a = synth.sphere_system(1200, 14.0, seed=5)
    b = synth.sphere_system(1200, 14.0, seed=5)
    for k in ("cmm", "q", "pol", "bonds_ri", "bonds_ci", "pgroup"):
        assert np.array_equal(a[k], b[k])
    deg = np.diff(a["bonds_ri"])
    assert deg.min() >= 1 and deg.max() <= 4
    cnt = np.bincount(a["pgroup"])
    assert cnt.min() >= 3 and cnt.max() <= 12
    w = synth.water_box()
    assert w["cmm"].shape == (648, 3)


def test_synthetic_groups_are_connected():
    from openmmpol_amd import synth
    a = synth.sphere_system(1500, 15.0, seed=9)
    ri, ci, g = a["bonds_ri"], a["bonds_ci"], a["pgroup"]
    for grp in range(g.max() + 1):
        mem = set(np.nonzero(g == grp)[0].tolist())
        start = next(iter(mem))
        seen, stack = {start}, [start]
        while stack:
            u = stack.pop()
            for v in ci[ri[u]:ri[u + 1]]:
                if v in mem and v not in seen:
                    seen.add(int(v))
                    stack.append(int(v))
        assert seen == mem


@pytest.mark.parametrize("name,ff", [("1crn", "AMOEBA"), ("1crn", "WANG_AL")])
def test_turn_pol_off_host_state_matches_oracle(name, ff, fixture_file):
    """ommp_turn_pol_off (mod_interface.F90:134-168): the polarizable set,
    positions, polarizabilities, Thole factors and screening lists afterwards
    against the oracle's restatement (tests/test_oracle.py pins that one)."""
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file(name, f"input_{ff}.mmp"))
    prep = mmp_io.prepare_mmp(mmp_io.fixture_path(name, f"input_{ff}.mmp"))
    o = oa.OracleSystem(prep)
    pm = o.polar_mm
    off = [int(pm[k]) + 1 for k in (0, 5, 6, 100, 101, 102, 300)] + [int(pm[-1]) + 1]
    s.turn_pol_off(off)
    o.turn_pol_off(off)
    assert s.pol_atoms == o.npol == len(pm) - len(off)
    assert np.array_equal(s.polar_mm, o.polar_mm)
    assert np.array_equal(s.cpol, prep["cmm"][o.polar_mm])
    assert np.array_equal(s.pol, prep["pol"][o.polar_mm])
    assert np.abs(s.thole - o.thole()).max() < 1e-15
    assert s.ipd.shape == (s.n_ipd, s.pol_atoms, 3) and not s.ipd.any()
    for which in range(0, 7):
        if which == 1 and not prep["amoeba"]:
            continue
        nrow = s.pol_atoms if which == 2 else s.mm_atoms
        assert _csr_dict(*s.get_list(which), nrow) == _csr_dict(*o.list_csr(which), nrow), which
    s.terminate()


def test_turn_pol_off_errors_exit_1_like_the_reference():
    """Indices outside 1..mm_atoms are fatal (mod_interface.F90:149-154)."""
    base = ("import openmmpol_amd as o, numpy as np\n"
            "s = o.OMMPSystem.from_arrays(True, np.zeros((2,3)) + [[0,0,0],[3,0,0]], np.zeros((2,10)), "
            "[1.0, 1.0], [0,0,0], [], pgroup=[0,1])\n")
    r = _run(base + "s.turn_pol_off([0])")
    assert r.returncode == 1 and "Atom indexes are 1-based, so no index should be less than 1" in r.stdout
    r = _run(base + "s.turn_pol_off([3])")
    assert r.returncode == 1 and "Atom indexes should be inside the topology range" in r.stdout
    r = _run(base + "s.turn_pol_off([1, 2]); print('npol', s.pol_atoms)")
    assert r.returncode == 0 and "npol 0" in r.stdout
