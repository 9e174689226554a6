"""MI355X (gfx950) induced-point-dipole polarization library with the
OpenMMPol C interface.  See include/openmmpol.h and include/ommp_amd.h."""
from . import pyopenmmpol
from .pyopenmmpol import (OMMPSystem, available_matrix_vector, available_solvers, set_verbose, time_pull,
                          time_push, verbosity)
from ._lib import LIB_PATH, lib
from .serving import Server

__all__ = ["OMMPSystem", "available_solvers", "available_matrix_vector", "verbosity", "set_verbose",
           "time_push", "time_pull", "lib", "LIB_PATH", "pyopenmmpol", "Server"]
