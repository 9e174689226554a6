"""Error behaviour at the C-ABI: a failure ends the process with the reference's
message format ("[OMMP]   [stop] ..." and exit status 1, src/mod_io.F90:155-175).

Each case runs in a child process (the library exits by design).  The failing
conditions are forced from the host -- an iteration limit too small to
converge, systems of two force fields in one server -- so every kernel runs to
its normal end.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PRELUDE = f"""
import sys
import numpy as np
sys.path.insert(0, {ROOT!r})
from openmmpol_amd import OMMPSystem, Server, synth
"""


def _run_child(body, timeout=120):
    code = PRELUDE + body
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=timeout,
                          cwd=ROOT)


@pytest.mark.parametrize("solver,precision,tol", [("cg", "fp64", 1e-14), ("diis", "fp64", 1e-14),
                                                  ("cg", "mixed", 1e-30)])
def test_not_converged_is_fatal(solver, precision, tol, fixture_file):
    """mod_solvers.F90:212-214 and :361-362: an iterative solve that does not
    reach the tolerance within its iteration limit is a fatal error (for the
    mixed solver: within its 30 refinement steps of 2 inner iterations each, the
    tolerance below what FP64 residuals can show)."""
    mmp = fixture_file("1crn", "input_AMOEBA.mmp")
    body = f"""
s = OMMPSystem({mmp!r})
s.set_precision({precision!r})
s.set_default_solver({solver!r})
s.set_solver_tolerance({tol!r}, 2)
print("solving", flush=True)
s.set_external_field(np.zeros((s.pol_atoms, 3)))
print("returned", flush=True)
"""
    r = _run_child(body)
    out = r.stdout + r.stderr
    print(out)
    assert r.returncode == 1, out
    assert "solving" in r.stdout and "returned" not in r.stdout
    assert "[OMMP]" in out and "[stop] Iterative solver did not converged" in out


def test_server_worker_failure_exits_from_caller(fixture_file):
    """A fatal error found on the server's worker thread (here: a second force
    field joining the batch) is reported by the caller's wait, which exits with
    the message; the worker never exits the process itself."""
    amoeba = fixture_file("N-methylacetamide", "input_AMOEBA.mmp")
    wang = fixture_file("N-methylacetamide", "input_WANG_AL.mmp")
    body = f"""
a = OMMPSystem({amoeba!r})
b = OMMPSystem({wang!r})
q = np.array([[[10.0, 0.0, 0.0, 0.5], [-10.0, 0.0, 0.0, -0.5]]] * 3)
srv = Server(hbm_budget=8e9, max_active=4)
srv.submit(a, q)
srv.submit(b, q)
print("waiting", flush=True)
srv.wait()
print("returned", flush=True)
"""
    r = _run_child(body)
    out = r.stdout + r.stderr
    print(out)
    assert r.returncode == 1, out
    assert "waiting" in r.stdout and "returned" not in r.stdout
    assert "[stop] All systems of a server must use the same force field" in out
