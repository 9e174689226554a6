"""GPU tests of the benchmark workloads not covered elsewhere.

648-site water box (216 rigid waters, one polarization group each, AMOEBA
default scales, zero external field, tolerance 1e-8): the plain FP64 CG against
the oracle's run of the same algorithm (dipoles 1e-9 a.u., energy 1e-10
relative, iteration counts within one), the mixed solver and the other
solvers against the oracle's exact solve (1e-7 a.u., 1e-9 relative).

60,000-site sphere: the mixed solver (on-the-fly and packed in-core FP32 T)
against the oracle's golden tight solve (tests/golden/make_sphere60k_golden.py):
energy 1e-9 relative, sampled dipoles 1e-7 a.u., per-set norms 1e-9 relative.
"""
import json
import os

import numpy as np
import pytest

import oracle_api as oa

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def water():
    from openmmpol_amd import synth
    spec = synth.water_box()
    return spec, oa.OracleSystem(synth.oracle_prep(spec))


def test_water_box_fp64_cg_follows_oracle(water):
    from openmmpol_amd import synth
    spec, o = water
    s = synth.make_system(spec)
    assert s.pol_atoms == 648
    ext = np.zeros((648, 3))
    s.set_external_field(ext)
    ep = s.get_polelec_energy()
    o.set_external_field(ext)
    eo = o.get_polelec_energy()
    assert np.abs(s.ipd - o.ipd).max() < 1e-9
    assert abs(ep - eo) <= 1e-10 * abs(eo)
    it = s.last_iterations
    assert all(abs(it[m] - o.iters[m]) <= 1 for m in range(2))
    s.terminate()


@pytest.mark.parametrize("mode,solver,matv", [("mixed", "cg", "direct"), ("mixed", "cg", "incore"),
                                              ("fp64", "diis", "direct"), ("fp64", "inversion", "direct")])
def test_water_box_solvers_match_exact_solve(water, mode, solver, matv):
    from openmmpol_amd import synth
    spec, o = water
    oa.lib().oracle_set_small_step_eps(0.0)
    try:
        want = o.polarization(o.m2d(), use_guess=False, tol=1e-12, maxit=1000)
    finally:
        oa.lib().oracle_set_small_step_eps(2.220446049250313080847e-16 * 100.0)
    ew = o.energy(want)
    s = synth.make_system(spec)
    s.set_precision(mode)
    s.set_solver_tolerance(1e-10, 500)
    s.set_default_solver(solver)
    s.set_default_matv(matv)
    s.set_external_field(np.zeros((648, 3)), nomm=False, solver=solver, matv=matv)
    ep = s.get_polelec_energy()
    assert np.abs(s.ipd - want).max() < 1e-7
    assert abs(ep - ew) <= 1e-9 * abs(ew)
    s.terminate()


@pytest.fixture(scope="module")
def sphere60k():
    path = os.path.join(GOLDEN, "sphere60000_seed3.json")
    if not os.path.exists(path):
        pytest.skip("60,000-site golden vectors not generated")
    from openmmpol_amd import synth
    with open(path) as f:
        g = json.load(f)
    spec = synth.sphere_system(g["n"], g["radius"], seed=g["seed"])
    return spec, g


@pytest.mark.parametrize("matv", ["direct", "incore"])
def test_sphere60k_mixed_matches_golden(sphere60k, matv):
    from openmmpol_amd import synth
    spec, g = sphere60k
    s = synth.make_system(spec)
    s.set_precision("mixed")
    s.set_solver_tolerance(3e-10, 500)
    s.set_default_matv(matv)
    em = s.e_m2d
    want_em = np.array(g["e_m2d_sample"])
    assert np.abs(em[:, g["sample"], :] - want_em).max() <= 1e-12 * np.abs(want_em).max()
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    ep = s.get_polelec_energy()
    ipd = s.ipd
    assert abs(ep - g["epol"]) <= 1e-9 * abs(g["epol"])
    assert np.abs(ipd[:, g["sample"], :] - np.array(g["ipd_sample"])).max() < 1e-7
    for k in range(2):
        assert abs(np.sqrt(np.sum(ipd[k] ** 2)) - g["ipd_norm"][k]) <= 1e-9 * g["ipd_norm"][k]
    s.terminate()
