"""Static field E_M2D (prepare_polelec, mod_electrostatics.F90:1487-1596) from
5,000 polarizable sites on, where each unordered pair of polarizable sites is
evaluated once (rhs_sym.hip) and the non-polarizable sources are swept
one-sided over their index list.  Bar: the oracle's E_M2D (its per-target
restatement of q/mu/quad_elec_prop with maxder-3 damping and the screening
lists) within 1e-12 of the largest component, the bar of the fixture systems
below that size (test_gpu_parity.py::test_static_field_matches_oracle).

Sizes leave a partial last block of 256 sites (7,001 = 27 x 256 + 89), and the
mixed system keeps every seventh site non-polarizable (Nmm 9,000, Npol 7,714)."""
import numpy as np
import pytest

import mmp_io
import oracle_api as oa

pytestmark = pytest.mark.gpu


def _radius(n):
    return (3.0 * n / (4.0 * np.pi * 0.0996)) ** (1.0 / 3.0)


def _check(s, o, label):
    e = np.asarray(s.e_m2d)
    eo = np.asarray(o.m2d())
    assert e.shape == eo.shape, (e.shape, eo.shape)
    err = np.abs(e - eo).max() / np.abs(eo).max()
    print(f"{label}: E_M2D err {err:.2e} of the largest component", flush=True)
    assert err <= 1e-12


def test_static_field_pair_once_amoeba():
    from openmmpol_amd import synth
    n = 7001
    spec = synth.sphere_system(n, _radius(n), seed=21)
    s = synth.make_system(spec)
    o = oa.OracleSystem(synth.oracle_prep(spec))
    _check(s, o, f"AMOEBA {n} sites")
    s.terminate()


def test_static_field_pair_once_wang():
    from openmmpol_amd import synth
    n = 7001
    spec = dict(synth.sphere_system(n, _radius(n), seed=22))
    spec["amoeba"] = False
    spec["q"] = spec["q"][:, :1].copy()
    spec["pgroup"] = None
    spec["scales"] = None  # the library's defaults for AMBER: Wang AL, as the oracle below
    s = synth.make_system(spec)
    prep = synth.oracle_prep(spec)
    prep["scales"] = mmp_io.SCALES_WANG_AL.copy()
    prep["thole_scale"] = mmp_io.THOLE_SCALE_WANG_AL
    o = oa.OracleSystem(prep)
    _check(s, o, f"Wang AL {n} sites")
    s.terminate()


def test_static_field_pair_once_with_nonpolarizable_sources():
    from openmmpol_amd import synth
    n = 9000
    spec = dict(synth.sphere_system(n, _radius(n), seed=23))
    spec["pol"] = spec["pol"].copy()
    spec["pol"][::7] = 0.0
    s = synth.make_system(spec)
    assert s.pol_atoms == n - len(range(0, n, 7)) and s.pol_atoms >= 5000
    o = oa.OracleSystem(synth.oracle_prep(spec))
    _check(s, o, f"AMOEBA {n} MM sites, {s.pol_atoms} polarizable")
    s.terminate()
