"""The inner matvec of the multi-RHS batched solver alone (ommp_amd_block_tmatvec):
damped tensors built in FP32, tensors and x split in three bf16 parts each, the
six leading products summed on the matrix cores with FP32 accumulation.

Bars, against the oracle's FP64 T x, for 1 to 128 columns, with P-P screening
(fixtures) and with sizes that leave partial source tiles and several source
splits (synthetic spheres): the FP32 kernels' bar (2e-6 of the largest
component, test_gpu_parity.py::test_tmatvec_kinds), and FP32-class accuracy:
within twice the error of the FP32 on-the-fly kernel on the same input, plus
5e-8 of the largest component.  Both kernels build the tensors in FP32, which
dominates their error (6e-8 to 6.4e-7 here).  A two-part split (hi + lo, three
products; about 2^-16 per product) fails the second bar on every input here:
emulated on the CPU for these inputs (the oracle's T rounded to FP32, products
summed exactly), it alone gives 0.98e-6 to 2.3e-6 of the largest component,
against 0.5e-8 to 1.4e-8 for the three-part split.
"""
import numpy as np
import pytest

import mmp_io
import oracle_api as oa

pytestmark = pytest.mark.gpu


def _check(s, o, ncols_list, seed):
    rng = np.random.default_rng(seed)
    nmax = max(ncols_list)
    x = rng.normal(0.0, 0.05, size=(nmax, s.pol_atoms, 3))
    ref = np.stack([o.tmatvec(x[k]) for k in range(nmax)])
    scale = np.abs(ref).max()
    y32, _ = s.tmatvec(x[:2], kind="otf32")
    err32 = np.abs(y32 - ref[:2]).max() / scale
    for ncols in ncols_list:
        y, _ = s.block_tmatvec(x[:ncols])
        err = np.abs(y - ref[:ncols]).max() / scale
        print(f"{s.pol_atoms} sites, {ncols} columns: block matvec err {err:.2e} of the largest "
              f"(FP32 on-the-fly kernel {err32:.2e})", flush=True)
        assert err <= 2e-6, ncols
        assert err <= 2.0 * err32 + 5e-8, ncols
        # no column is skipped or mixed with another
        assert np.all(np.abs(y).max(axis=(1, 2)) > 0.0)


@pytest.mark.parametrize("name,ff", [("1crn", "AMOEBA"), ("1crn", "WANG_AL"), ("N-methylacetamide", "AMOEBA")])
def test_block_matvec_matches_oracle_fixtures(name, ff, fixture_file):
    from openmmpol_amd import OMMPSystem
    s = OMMPSystem(fixture_file(name, f"input_{ff}.mmp"))
    o = oa.OracleSystem(mmp_io.prepare_mmp(mmp_io.fixture_path(name, f"input_{ff}.mmp")))
    _check(s, o, (1, 2, 7, 128), 5)
    s.terminate()


@pytest.mark.parametrize("n", [3000, 4001])
def test_block_matvec_matches_oracle_sphere(n):
    """Synthetic AMOEBA spheres: 3,000 sites (a multiple of the 32-site tiles)
    and 4,001 (a partial last tile in the last split)."""
    from openmmpol_amd import synth
    spec = synth.sphere_system(n, (3.0 * n / (4.0 * np.pi * 0.0996)) ** (1.0 / 3.0), seed=13)
    s = synth.make_system(spec)
    o = oa.OracleSystem(synth.oracle_prep(spec))
    _check(s, o, (3, 16), 6)
    s.terminate()
