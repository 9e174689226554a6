"""GPU tests of the batched solver: multi-RHS solves and the serving loop.

Bars: every column against the oracle's tight FP64 solve of the same field
(dipoles 1e-7 a.u. max-abs, E_pol 1e-9 relative, solver tolerance 1e-10 on the
preconditioned rms); the point-charge field against the oracle to 1e-12
relative; a request's results alone and in a mix of other requests identical
(bit for bit: a column's arithmetic does not depend on the rest of the batch).
"""
import math

import numpy as np
import pytest

import mmp_io
import oracle_api as oa

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture
def tight_oracle():
    """Oracle with the reference's |p.Tp| quick return disabled (exact solves)."""
    oa.lib().oracle_set_small_step_eps(0.0)
    yield
    oa.lib().oracle_set_small_step_eps(2.220446049250313080847e-16 * 100.0)


def _exact(o, ext):
    E = o.m2d()
    ipd = o.polarization(E + ext[None], use_guess=False, tol=1e-12, maxit=2000)
    return ipd, o.energy(ipd)


def _qm(rng, nq=30, cavity=6.0):
    from openmmpol_amd import synth
    cq, qq = synth.qm_charges(rng, nq, cavity=cavity)
    return np.concatenate([cq, qq[:, None]], axis=1)


@pytest.mark.parametrize("ff", ["AMOEBA", "WANG_AL"])
def test_charges_field_matches_oracle(ff, fixture_file):
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file("1crn", f"input_{ff}.mmp"))
    o = oa.OracleSystem(mmp_io.prepare_mmp(mmp_io.fixture_path("1crn", f"input_{ff}.mmp")))
    c = np.asarray(s.cmm)[np.asarray(s.polar_mm)]
    rng = np.random.default_rng(5)
    q = np.concatenate([c.mean(0) + rng.normal(0, 4.0, size=(12, 3)), rng.normal(0, 0.5, size=(12, 1))], axis=1)
    got = s.charges_field(q)
    want = o.charges_field(np.ascontiguousarray(q[:, :3]), np.ascontiguousarray(q[:, 3]))
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    s.terminate()


@pytest.mark.parametrize("ff", ["AMOEBA", "WANG_AL"])
@pytest.mark.parametrize("nrhs", [1, 5])
def test_multi_rhs_matches_oracle(ff, nrhs, fixture_file, tight_oracle):
    from openmmpol_amd import OMMPSystem
    name = "1crn"
    s = OMMPSystem(fixture_file(name, f"input_{ff}.mmp"))
    s.set_solver_tolerance(TOL, 500)
    o = oa.OracleSystem(mmp_io.prepare_mmp(mmp_io.fixture_path(name, f"input_{ff}.mmp")))
    ef1 = mmp_io.read_ef(mmp_io.fixture_path(name, "EF_1.txt"))[np.asarray(s.polar_mm)]
    rng = np.random.default_rng(3)
    fields = [ef1 * (0.5 + 0.25 * k) + 1e-3 * rng.normal(size=ef1.shape) for k in range(nrhs)]
    ext = np.stack(fields)
    ipd, ep = s.set_external_fields_multi(ext)
    assert ipd.shape == (nrhs, s.n_ipd, s.pol_atoms, 3)
    for k in range(nrhs):
        want, ew = _exact(o, ext[k])
        assert np.abs(ipd[k] - want).max() < 1e-7, k
        assert abs(ep[k] - ew) <= 1e-9 * abs(ew), k
    # the system's own state is untouched: a normal solve afterwards still works
    s.set_external_field(ext[0])
    assert np.abs(s.ipd - ipd[0]).max() < 1e-7
    s.terminate()


def test_multi_rhs_warm_start_and_repeat(fixture_file, tight_oracle):
    from openmmpol_amd import OMMPSystem
    name, ff = "1crn", "AMOEBA"
    s = OMMPSystem(fixture_file(name, f"input_{ff}.mmp"))
    s.set_solver_tolerance(TOL, 500)
    o = oa.OracleSystem(mmp_io.prepare_mmp(mmp_io.fixture_path(name, f"input_{ff}.mmp")))
    ef1 = mmp_io.read_ef(mmp_io.fixture_path(name, "EF_1.txt"))[np.asarray(s.polar_mm)]
    ext = np.stack([ef1 * (1.0 + 0.01 * k) for k in range(9)])
    ipd_cold, ep_cold = s.set_external_fields_multi(ext)
    cold_iters = s.last_iterations[3]
    s.set_external_field(ef1)  # leaves a guess
    ipd_warm, ep_warm = s.set_external_fields_multi(ext)
    assert s.last_iterations[3] < cold_iters
    for k in (0, 4, 8):
        want, ew = _exact(o, ext[k])
        assert np.abs(ipd_warm[k] - want).max() < 1e-7
        assert abs(ep_warm[k] - ew) <= 1e-9 * abs(ew)
    # a second call with the same guess reuses the cached batch state exactly
    ipd_again, ep_again = s.set_external_fields_multi(ext)
    assert np.array_equal(ipd_again, ipd_warm) and np.array_equal(ep_again, ep_warm)
    s.terminate()


# ------------------------------------------------------------------ serving
SIZES = [(300, 11), (700, 12), (1500, 13), (2600, 14)]


def _sphere(n, seed):
    from openmmpol_amd import synth
    radius = (3.0 * n / (4.0 * math.pi * 0.1)) ** (1.0 / 3.0) + 6.0 * 0.5
    return synth.sphere_system(n, radius, seed=seed, cavity=6.0)


def _charges(seed, nsteps):
    rng = np.random.default_rng(seed)
    base = _qm(rng)
    steps = []
    for _ in range(nsteps):
        q = base.copy()
        q[:, :3] += rng.normal(0.0, 0.02 * 1.8897261245650, size=(len(q), 3))
        steps.append(q)
    return np.stack(steps)


@pytest.fixture(scope="module")
def requests():
    specs = [_sphere(n, sd) for n, sd in SIZES]
    charges = [_charges(100 + k, 3 + k % 2) for k in range(len(specs))]
    return specs, charges


def _run(specs, charges, which, budget=64e9, max_active=8):
    from openmmpol_amd import Server, synth
    systems = {k: synth.make_system(specs[k]) for k in which}
    for s in systems.values():
        s.set_solver_tolerance(TOL, 500)
    srv = Server(hbm_budget=budget, max_active=max_active)
    ids = {k: srv.submit(systems[k], charges[k]) for k in which}
    srv.wait()
    out = {}
    for k in which:
        st, e, lat, it = srv.result(ids[k])
        assert st == 0
        assert np.all(lat > 0) and np.all(it[:, 1] > 0)
        out[k] = (e, systems[k].ipd.copy())
    stats = srv.stats()
    srv.close()
    for s in systems.values():
        s.terminate()
    return out, stats


def test_server_alone_and_in_mix_agree(requests, tight_oracle):
    specs, charges = requests
    alone = {}
    for k in range(len(specs)):
        r, _ = _run(specs, charges, [k])
        alone[k] = r[k]
    mix, stats = _run(specs, charges, list(range(len(specs))))
    assert stats["max_admitted"] >= 2
    for k in range(len(specs)):
        assert np.array_equal(alone[k][0], mix[k][0]), k
        assert np.array_equal(alone[k][1], mix[k][1]), k
    # every solve against the oracle's exact solve of its field
    from openmmpol_amd import synth
    for k in (0, 2):
        o = oa.OracleSystem(synth.oracle_prep(specs[k]))
        cpol = specs[k]["cmm"][o.polar_mm]
        for step in range(len(charges[k])):
            q = charges[k][step]
            ext = synth.point_charge_field(cpol, q[:, :3], q[:, 3])
            ipd, ew = _exact(o, ext)
            assert abs(mix[k][0][step] - ew) <= 1e-9 * abs(ew), (k, step)
        assert np.abs(mix[k][1] - ipd).max() < 1e-7


def test_server_admission_waits_for_memory(requests):
    specs, charges = requests
    which = [0, 1, 2]
    mix, _ = _run(specs, charges, which)
    # a budget below any system's footprint: each request waits until the
    # previous one has released its memory (one admitted at a time)
    tight, stats = _run(specs, charges, which, budget=1.0)
    assert stats["deferrals"] > 0 and stats["max_admitted"] == 1
    for k in which:
        assert np.array_equal(mix[k][0], tight[k][0]) and np.array_equal(mix[k][1], tight[k][1])


# ------------------------------------------------ pair-once sweep in batches
def test_multi_rhs_pair_once_matches_oracle(tight_oracle):
    """A 7,500-site system (30 blocks of 256, the last one padded): its FP64
    refinement matvec in the batch runs the pair-once sweep, two columns per
    pass over a group of four; every field against the oracle's exact solve."""
    from openmmpol_amd import synth
    spec = _sphere(7500, 31)
    s = synth.make_system(spec)
    s.set_solver_tolerance(TOL, 500)
    o = oa.OracleSystem(synth.oracle_prep(spec))
    cpol = spec["cmm"][o.polar_mm]
    rng = np.random.default_rng(9)
    ext = []
    for k in range(3):
        q = _qm(rng)
        ext.append(synth.point_charge_field(cpol, q[:, :3], q[:, 3]))
    ext = np.stack(ext)
    ipd, ep = s.set_external_fields_multi(ext)
    for k in range(len(ext)):
        want, ew = _exact(o, ext[k])
        assert np.abs(ipd[k] - want).max() < 1e-7, k
        assert abs(ep[k] - ew) <= 1e-9 * abs(ew), k
    s.terminate()


def test_server_pair_once_sizes_alone_and_in_mix(tight_oracle):
    """Requests of 7,500 sites (FP64 matvec pair-once), 13,000 (both matvecs
    pair-once) and 700 (per-target) together: each request's results alone and
    in the mix are identical, and the 7,500-site one matches the oracle."""
    specs = [_sphere(n, sd) for n, sd in ((7500, 41), (13000, 42), (700, 43))]
    charges = [_charges(200 + k, 2) for k in range(len(specs))]
    alone = {}
    for k in range(len(specs)):
        r, _ = _run(specs, charges, [k])
        alone[k] = r[k]
    mix, stats = _run(specs, charges, list(range(len(specs))))
    assert stats["max_admitted"] >= 2
    for k in range(len(specs)):
        assert np.array_equal(alone[k][0], mix[k][0]), k
        assert np.array_equal(alone[k][1], mix[k][1]), k
    from openmmpol_amd import synth
    o = oa.OracleSystem(synth.oracle_prep(specs[0]))
    cpol = specs[0]["cmm"][o.polar_mm]
    for step in range(len(charges[0])):
        q = charges[0][step]
        ext = synth.point_charge_field(cpol, q[:, :3], q[:, 3])
        ipd, ew = _exact(o, ext)
        assert abs(mix[0][0][step] - ew) <= 1e-9 * abs(ew), step
    assert np.abs(mix[0][1] - ipd).max() < 1e-7


def test_server_serving_workload_matches_oracle_golden():
    """The largest request of the serving workload (39,855 sites) together with
    a 633-site one, through the server, against the oracle's tight solves of
    their first and last fields (tests/golden/make_serving_golden.py)."""
    import json
    import os
    from openmmpol_amd import Server, synth
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "serving16_seed5.json")) as f:
        g = json.load(f)
    sizes, nsteps, _, charges, make_spec = synth.serving_requests(g["nreq"], g["seed"], g["mean_gap_ms"])
    which = (9, 4)
    systems = {k: synth.make_system(make_spec(k)) for k in which}
    for s in systems.values():
        s.set_solver_tolerance(3e-10, 500)
    srv = Server(hbm_budget=200e9, max_active=16)
    ids = {k: srv.submit(systems[k], charges[k]) for k in which}
    srv.wait()
    cases = [c for c in g["cases"] if c["request"] in which]
    assert {c["request"] for c in cases} == set(which)
    for k in which:
        st, e, _, _ = srv.result(ids[k])
        assert st == 0
        ipd = systems[k].ipd
        for c in (c for c in cases if c["request"] == k):
            print(f"request {k} ({sizes[k]} sites) step {c['step']}: E_pol {e[c['step']]:.12e} "
                  f"oracle {c['epol']:.12e}", flush=True)
            assert abs(e[c["step"]] - c["epol"]) <= 1e-9 * abs(c["epol"]), (k, c["step"])
            if c["step"] == nsteps[k] - 1:
                assert np.abs(ipd[:, c["sample"], :] - np.array(c["ipd_sample"])).max() <= 1e-7, k
    srv.close()
    for s in systems.values():
        s.terminate()
