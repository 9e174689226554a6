"""15,000-site synthetic AMOEBA system against the oracle's golden vectors
(tests/golden/make_sphere_golden.py: FP64 CG to 1e-12).  Bars: polarization
energy within 1e-9 relative and every induced dipole within 1e-7 a.u."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def sphere():
    from openmmpol_amd import synth
    with open(os.path.join(GOLDEN, "sphere15000_seed2.json")) as f:
        g = json.load(f)
    ref = np.load(os.path.join(GOLDEN, "sphere15000_seed2_ipd_zero.npy"))
    spec = synth.sphere_system(g["n"], g["radius"], seed=g["seed"])
    return spec, g, ref


def test_static_field_matches_golden(sphere):
    from openmmpol_amd import synth
    spec, g, _ = sphere
    s = synth.make_system(spec)
    em = s.e_m2d
    want = np.array(g["e_m2d_sample"])
    assert np.abs(em[:, g["sample"], :] - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("matv", ["direct", "incore"])
def test_mixed_solver_matches_golden(sphere, matv):
    from openmmpol_amd import synth
    spec, g, ref = sphere
    s = synth.make_system(spec)
    s.set_precision("mixed")
    s.set_default_matv(matv)
    s.set_solver_tolerance(3e-10, 500)
    s.set_external_field(np.zeros((s.pol_atoms, 3)))
    e = s.get_polelec_energy()
    eref = g["cases"]["zero"]["epol"]
    assert abs(e - eref) <= 1e-9 * abs(eref)
    assert np.abs(s.ipd - ref).max() <= 1e-7


def test_fp64_solver_matches_golden_qm_field(sphere):
    """FP64 CG with a 30-point-charge field (the energy getter excludes the
    external field, mod_electrostatics.F90:963-974)."""
    from openmmpol_amd import synth
    spec, g, _ = sphere
    s = synth.make_system(spec)
    cq = np.array(g["qm_charges"]["pos_bohr"])
    qq = np.array(g["qm_charges"]["q"])
    ext = synth.point_charge_field(s.cpol, cq, qq)
    s.set_solver_tolerance(5e-10, 500)
    s.set_external_field(ext)
    e = s.get_polelec_energy()
    eref = g["cases"]["qm30"]["epol"]
    assert abs(e - eref) <= 1e-9 * abs(eref)
    ipd = s.ipd
    assert np.abs(ipd[:, g["sample"], :] - np.array(g["cases"]["qm30"]["ipd_sample"])).max() <= 1e-7


def test_warm_started_scf_sequence(sphere):
    """Solves with slowly changing fields reuse the previous dipoles as guess and
    keep the accuracy of a cold solve."""
    from openmmpol_amd import synth
    spec, g, _ = sphere
    s = synth.make_system(spec)
    s.set_precision("mixed")
    s.set_solver_tolerance(3e-10, 500)
    cq = np.array(g["qm_charges"]["pos_bohr"])
    qq = np.array(g["qm_charges"]["q"])
    rng = np.random.default_rng(3)
    inner = []
    for step in range(4):
        ext = synth.point_charge_field(s.cpol, cq + rng.normal(0, 0.02 * 1.8897261245650, cq.shape), qq)
        s.set_external_field(ext)
        inner.append(s.last_iterations[3])
        e_warm = s.get_polelec_energy()
        ipd_warm = s.ipd.copy()
    s.reset_guess()
    s.set_external_field(ext)
    assert np.abs(s.ipd - ipd_warm).max() <= 1e-7
    assert abs(s.get_polelec_energy() - e_warm) <= 1e-9 * abs(e_warm)
    assert inner[-1] < inner[0]


def test_mixed_solves_enqueued_ahead(sphere):
    """From the second mixed solve on, the refinement steps are enqueued without
    host round trips with the previous solve's step and iteration counts, every
    decision taken on the device.  Repeated cold solves give the same dipoles bit
    for bit; a tighter tolerance than the prediction (more steps: the host-driven
    steps take over) and a looser one (fewer steps: the rest are skipped) stay
    at the accuracy of a fresh solve."""
    from openmmpol_amd import synth
    spec, g, ref = sphere
    eref = g["cases"]["zero"]["epol"]
    s = synth.make_system(spec)
    s.set_precision("mixed")
    s.set_solver_tolerance(3e-10, 500)
    zero = np.zeros((s.pol_atoms, 3))
    runs = []
    for k in range(4):
        s.reset_guess(True)
        s.set_external_field(zero)
        runs.append((s.get_polelec_energy(), s.ipd.copy(), list(s.last_iterations)))
        print(f"cold solve {k}: iterations {runs[-1][2]}", flush=True)
    for e, ipd, _ in runs:
        assert abs(e - eref) <= 1e-9 * abs(eref)
        assert np.abs(ipd - ref).max() <= 1e-7
    for e, ipd, it in runs[2:]:
        assert e == runs[1][0] and np.array_equal(ipd, runs[1][1]) and it == runs[1][2]
    for tol in (1e-11, 3e-10, 1e-8, 3e-10):
        s.set_solver_tolerance(tol, 500)
        s.reset_guess(True)
        s.set_external_field(zero)
        e = s.get_polelec_energy()
        print(f"tol {tol}: iterations {list(s.last_iterations)}", flush=True)
        bar = 1e-9 if tol < 1e-9 else 1e-7
        assert abs(e - eref) <= bar * abs(eref)
        assert np.abs(s.ipd - ref).max() <= (1e-7 if tol < 1e-9 else 1e-5)
    # a warm-started solve after cold ones and an FP64 CG solve (which leaves its
    # own work vectors behind): a small field change needs one inner solve
    # between two FP64 residuals; then larger changes, against a fresh system
    cq = np.array(g["qm_charges"]["pos_bohr"])
    qq = np.array(g["qm_charges"]["q"])
    ext = synth.point_charge_field(s.cpol, cq, qq)
    s.set_precision("fp64")
    s.set_external_field(2e-3 * ext)
    s.set_precision("mixed")
    s.set_external_field(1e-3 * ext)
    print(f"warm solve after cold ones: iterations {list(s.last_iterations)}", flush=True)
    assert s.last_iterations[2] <= 2
    for _ in range(3):
        s.set_external_field(ext)
    e = s.get_polelec_energy()
    f = synth.make_system(spec)
    f.set_precision("mixed")
    f.set_solver_tolerance(3e-10, 500)
    f.set_external_field(ext)
    assert abs(e - f.get_polelec_energy()) <= 1e-9 * abs(e)
    assert np.abs(s.ipd - f.ipd).max() <= 1e-7
    s.terminate()
    f.terminate()


def test_multi_rhs_64_fields_matches_golden():
    """64 QM-region fields on the 15,000-site cavity system in one batched call
    (matrix-core FP32 block matvec, FP64 refinement per column) against the
    oracle's tight solves of 19 of them spread over the batch, the last column
    group included (tests/golden/make_multi_golden.py)."""
    from openmmpol_amd import synth
    with open(os.path.join(GOLDEN, "multi15000_cavity.json")) as f:
        g = json.load(f)
    ref0 = np.load(os.path.join(GOLDEN, "multi15000_cavity_ipd0.npy"))
    spec = synth.sphere_system(g["n"], g["radius"], seed=g["seed"], cavity=g["cavity"])
    charges = synth.qm_charge_sets(g["nrhs"], g["qm_seed"])
    s = synth.make_system(spec)
    s.set_solver_tolerance(3e-10, 500)
    cpol = spec["cmm"][np.asarray(s.polar_mm)]
    ext = np.stack([synth.point_charge_field(cpol, q[:, :3], q[:, 3]) for q in charges])
    ipd, ep = s.set_external_fields_multi(ext)
    fields = [r["field"] for r in g["rhs"]]
    assert len(fields) >= 16 and fields[-1] == g["nrhs"] - 1
    for r in g["rhs"]:
        k = r["field"]
        assert abs(ep[k] - r["epol"]) <= 1e-9 * abs(r["epol"]), k
        assert np.abs(ipd[k][:, g["sample"], :] - np.array(r["ipd_sample"])).max() <= 1e-7, k
    assert np.abs(ipd[0] - ref0).max() <= 1e-7
    s.terminate()
