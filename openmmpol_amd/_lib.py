"""ctypes bindings over libopenmmpol_amd.so (the C-ABI declared in include/).

The library is the only compute path: there is no CPU fallback.  Loading fails
loudly when the shared object has not been built.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopenmmpol_amd.so")

_lib = None

c_dp = ctypes.POINTER(ctypes.c_double)
c_ip = ctypes.POINTER(ctypes.c_int32)
c_vp = ctypes.c_void_p

# (name, restype, argtypes) for every entry point of include/openmmpol.h and include/ommp_amd.h
SIGNATURES = [
    ("ommp_init_mmp", c_vp, [ctypes.c_char_p]),
    ("ommp_set_default_solver", None, [c_vp, ctypes.c_int32]),
    ("ommp_set_default_matv", None, [c_vp, ctypes.c_int32]),
    ("ommp_terminate", None, [c_vp]),
    ("ommp_set_verbose", None, [ctypes.c_int32]),
    ("ommp_set_outputfile", None, [ctypes.c_char_p]),
    ("ommp_close_outputfile", None, []),
    ("ommp_message", None, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p]),
    ("ommp_fatal", None, [ctypes.c_char_p]),
    ("ommp_time_pull", None, [ctypes.c_char_p]),
    ("ommp_time_push", None, []),
    ("ommp_get_polelec_energy", ctypes.c_double, [c_vp]),
    ("ommp_get_fixedelec_energy", ctypes.c_double, [c_vp]),
    ("ommp_get_full_ele_energy", ctypes.c_double, [c_vp]),
    ("ommp_set_external_field", None, [c_vp, c_dp, ctypes.c_int32, ctypes.c_int32]),
    ("ommp_set_external_field_nomm", None, [c_vp, c_dp, ctypes.c_int32, ctypes.c_int32]),
    ("ommp_ff_is_amoeba", ctypes.c_bool, [c_vp]),
    ("ommp_get_n_ipd", ctypes.c_int32, [c_vp]),
    ("ommp_get_ld_cart", ctypes.c_int32, [c_vp]),
    ("ommp_get_mm_atoms", ctypes.c_int32, [c_vp]),
    ("ommp_get_pol_atoms", ctypes.c_int32, [c_vp]),
    ("ommp_get_cmm", c_dp, [c_vp]),
    ("ommp_get_cpol", c_dp, [c_vp]),
    ("ommp_get_q", c_dp, [c_vp]),
    ("ommp_get_ipd", c_dp, [c_vp]),
    ("ommp_get_polar_mm", c_ip, [c_vp]),
    ("ommp_get_zmm", c_ip, [c_vp]),
    ("ommp_update_coordinates", None, [c_vp, c_dp]),
    ("ommp_turn_pol_off", None, [c_vp, ctypes.c_int32, c_ip]),
    ("ommp_enable_fmm", None, [c_vp]),
    ("ommp_disable_fmm", None, [c_vp]),
    ("ommp_use_fmm", ctypes.c_bool, [c_vp]),
    ("ommp_set_fmm_lmax_pol", None, [c_vp, ctypes.c_int32]),
    ("ommp_set_fmm_lmax", None, [c_vp, ctypes.c_int32]),
    ("ommp_set_fmm_distance", None, [c_vp, ctypes.c_double]),
    ("ommp_set_fmm_min_cell_size", None, [c_vp, ctypes.c_double]),
    ("ommp_amd_init_arrays", c_vp, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, c_dp, c_dp, c_dp, c_dp,
                                    c_ip, c_ip, c_ip, c_ip]),
    ("ommp_amd_set_screening_parameters", None, [c_vp, c_dp, c_dp, c_dp, c_dp, c_dp]),
    ("ommp_amd_set_solver_tolerance", None, [c_vp, ctypes.c_double, ctypes.c_int32]),
    ("ommp_amd_set_precision", None, [c_vp, ctypes.c_int32]),
    ("ommp_amd_get_precision", ctypes.c_int32, [c_vp]),
    ("ommp_amd_get_last_iterations", None, [c_vp, c_ip]),
    ("ommp_amd_get_last_status", ctypes.c_int32, [c_vp]),
    ("ommp_amd_get_e_m2d", c_dp, [c_vp]),
    ("ommp_amd_set_external_field_device", None, [c_vp, c_vp, ctypes.c_int32]),
    ("ommp_amd_device_ipd", c_vp, [c_vp]),
    ("ommp_amd_set_stream", None, [c_vp, c_vp]),
    ("ommp_amd_reset_guess", None, [c_vp, ctypes.c_bool]),
    ("ommp_amd_set_inner_preconditioner", None, [c_vp, ctypes.c_int32]),
    ("ommp_amd_inner_precond_apply", ctypes.c_int32, [c_vp, ctypes.c_int32, c_dp, c_dp]),
    ("ommp_amd_get_list", ctypes.c_int32, [c_vp, ctypes.c_int32, c_ip, c_ip, c_dp]),
    ("ommp_amd_get_thole", c_dp, [c_vp]),
    ("ommp_amd_get_pol", c_dp, [c_vp]),
    ("ommp_amd_get_last_timings", None, [c_vp, c_dp]),
    ("ommp_amd_vector_fma_tflops", ctypes.c_double, [ctypes.c_int32]),
    ("ommp_amd_profile_kernels", None, [c_vp, ctypes.c_bool]),
    ("ommp_amd_profile_kernels_every", None, [c_vp, ctypes.c_int32]),
    ("ommp_amd_get_kernel_stats", None, [c_vp, c_dp, ctypes.c_bool]),
    ("ommp_amd_tmatvec", ctypes.c_double, [c_vp, ctypes.c_int32, ctypes.c_int32, c_dp, c_dp, ctypes.c_int32]),
    ("ommp_amd_block_tmatvec", ctypes.c_double, [c_vp, ctypes.c_int32, c_dp, c_dp, ctypes.c_int32]),
    ("ommp_potential_mmpol2ext", None, [c_vp, ctypes.c_int32, c_dp, c_dp]),
    ("ommp_potential_mm2ext", None, [c_vp, ctypes.c_int32, c_dp, c_dp]),
    ("ommp_potential_pol2ext", None, [c_vp, ctypes.c_int32, c_dp, c_dp]),
    ("ommp_field_mmpol2ext", None, [c_vp, ctypes.c_int32, c_dp, c_dp]),
    ("ommp_amd_field_mmpol2ext_noalloc", None, [c_vp, ctypes.c_int32, c_dp, c_dp]),
    ("ommp_field_mm2ext", None, [c_vp, ctypes.c_int32, c_dp, c_dp]),
    ("ommp_field_pol2ext", None, [c_vp, ctypes.c_int32, c_dp, c_dp]),
    ("ommp_polelec_geomgrad", None, [c_vp, c_dp]),
    ("ommp_rotation_geomgrad", None, [c_vp, c_dp, c_dp, c_dp]),
    ("ommp_amd_set_external_fields_multi", None, [c_vp, ctypes.c_int32, c_dp, c_dp, c_dp]),
    ("ommp_amd_set_external_fields_multi_device", None, [c_vp, ctypes.c_int32, c_vp, c_vp, c_dp]),
    ("ommp_amd_charges_field", None, [c_vp, ctypes.c_int32, c_dp, c_dp]),
    ("ommp_amd_server_create", c_vp, [ctypes.c_double, ctypes.c_int32]),
    ("ommp_amd_server_submit", ctypes.c_int32, [c_vp, c_vp, ctypes.c_int32, ctypes.c_int32, c_dp]),
    ("ommp_amd_server_wait", None, [c_vp, ctypes.c_int32]),
    ("ommp_amd_server_result", ctypes.c_int32, [c_vp, ctypes.c_int32, c_dp, c_dp, c_ip]),
    ("ommp_amd_server_stats", None, [c_vp, c_dp]),
    ("ommp_amd_server_destroy", None, [c_vp]),
]


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C openmmpol_amd/csrc`; there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
