"""The no-allocation entry point ommp_amd_field_mmpol2ext_noalloc against the
allocating ommp_field_mmpol2ext it mirrors, through the C-ABI: the same
kernels and split counts, so the same bits; two calls in a row, the second
with fewer points (the shared scratch array keeps the first call's contents),
on both force fields.  The size limit is checked in a child process (the
library exits on it, as on every error)."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import mmp_io

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _field(L, h, fn, cext, e0):
    e = np.ascontiguousarray(e0.copy())
    c = np.ascontiguousarray(cext)
    getattr(L, fn)(h, ctypes.c_int32(len(c)), c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                   e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return e


@pytest.mark.parametrize("ff", ["AMOEBA", "WANG_AL"])
def test_noalloc_field_equals_allocating_entry(ff, fixture_file):
    from openmmpol_amd import OMMPSystem, lib
    L = lib()
    name = "1crn"
    s = OMMPSystem(fixture_file(name, f"input_{ff}.mmp"))
    ef = mmp_io.read_ef(mmp_io.fixture_path(name, "EF_1.txt"))[np.asarray(s.polar_mm)]
    s.set_external_field(ef)  # dipoles for the induced part; the device state exists from here on
    rng = np.random.default_rng(11)
    c = np.asarray(s.cmm)
    lo, hi = c.min(0) - 8.0, c.max(0) + 8.0
    for n in (8192, 300):  # the largest call, then a smaller one over the same scratch
        cext = rng.uniform(lo, hi, size=(n, 3))
        e0 = rng.normal(0.0, 1e-3, size=(n, 3))  # results are added to the caller's array
        want = _field(L, s.handler, "ommp_field_mmpol2ext", cext, e0)
        got = _field(L, s.handler, "ommp_amd_field_mmpol2ext_noalloc", cext, e0)
        print(f"{ff} n={n}: max |noalloc - alloc| = {np.abs(got - want).max():.3e}", flush=True)
        assert np.array_equal(got, want), n
    s.terminate()


def test_noalloc_field_rejects_too_many_points(fixture_file):
    mmp = fixture_file("N-methylacetamide", "input_AMOEBA.mmp")
    code = f"""
import sys
import numpy as np
sys.path.insert(0, {ROOT!r})
import ctypes
from openmmpol_amd import OMMPSystem, lib
s = OMMPSystem({mmp!r})
s.set_external_field(np.zeros((s.pol_atoms, 3)))
n = 8193
c = np.full((n, 3), 50.0)
e = np.zeros((n, 3))
print("calling", flush=True)
lib().ommp_amd_field_mmpol2ext_noalloc(s.handler, n, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
print("returned", flush=True)
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, cwd=ROOT)
    out = r.stdout + r.stderr
    print(out)
    assert r.returncode == 1 and "calling" in r.stdout and "returned" not in r.stdout, out
    assert "takes at most 8192 external points per call (8193 requested)" in out
