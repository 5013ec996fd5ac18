"""ctypes binding of libnegf_amd.so (the HIP product library).

The library is loaded from electrostatic_amd/lib/; there is no CPU fallback:
if it is missing, loading raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnegf_amd.so")

NEGF_OK = 0
STATUS_NAMES = {0: "NEGF_OK", 1: "NEGF_EARG", 2: "NEGF_EALLOC", 3: "NEGF_EROOT",
                4: "NEGF_ENONCONV", 5: "NEGF_ENAN", 6: "NEGF_EDEVICE",
                7: "NEGF_ENOTSUP"}


class NegfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class negf_complex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class negf_chain_params(C.Structure):
    _fields_ = [("N", C.c_int), ("t_even", C.c_double), ("t_odd", C.c_double),
                ("Uc", C.c_double * 2), ("use_decimation", C.c_int),
                ("decimation_layers", C.c_int),
                ("decimation_max_iter", C.c_int),
                ("decimation_rel_error", C.c_int),
                ("segment_length", C.c_int)]


class negf_tube_params(C.Structure):
    _fields_ = [
        ("type_id", C.c_int * 2),
        ("acc", C.c_double),
        ("gamma", C.c_double),
        ("num_modes", C.c_int),
        ("num_unitcells", C.c_int),
        ("E_valence_min", C.c_double),
        ("E_pole_max", C.c_double),
        ("E_zPlus_imag", C.c_double),
        ("Fermi_tail_factors", C.c_double * 2),
        ("eq_integration_pts", C.c_int * 3),
        ("num_noneq_paths", C.c_int),
        ("noneq_integration_pts", C.c_int * 8),
        ("noneq_percent_intercuts", C.c_double * 8),
        ("contact_mu", C.c_double * 2),
        ("contact_T", C.c_double * 2),
        ("use_decimation", C.c_int),
        ("decimation_layers", C.c_int),
        ("decimation_max_iter", C.c_int),
        ("decimation_rel_error", C.c_int),
        ("flag_adaptive_integration_limits", C.c_int),
        ("integrand_correction_interval", C.c_int),
        ("kT_window_around_singularity", C.c_double * 2),
        ("noneq_integration_pts_density", C.c_int * 8),
        ("flatband_dos_integration_pts", C.c_int),
        ("flatband_dos_integration_limits", C.c_double * 2),
        ("segment_length", C.c_int),
    ]


class negf_scf_params(C.Structure):
    _fields_ = [("eps_r", C.c_double), ("R_ring_nm", C.c_double),
                ("R_image_nm", C.c_double), ("Broyden_fraction", C.c_double),
                ("Broyden_max_norm", C.c_double),
                ("Broyden_norm_type", C.c_int),
                ("Broyden_threshold_maxstep", C.c_int),
                ("reset_with_previous_charge_distribution", C.c_int),
                ("NS_initial_deposit_value", C.c_double)]


dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "negf_last_error": (C.c_char_p, []),
    "negf_version": (C.c_int, []),
    "negf_max_sites": (C.c_int, []),
    "negf_multi_scf_create": (C.c_int, [vp, vp, C.c_int, vp, vp]),
    "negf_multi_scf_destroy": (C.c_int, [vp]),
    "negf_multi_scf_num_sites": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
    "negf_multi_scf_set_external_potential": (C.c_int, [vp, vp]),
    "negf_multi_scf_use_graph": (C.c_int, [vp, C.c_int]),
    "negf_multi_scf_bias_point": (C.c_int, [vp, vp, C.c_int, vp, vp, vp, vp, vp]),
    "negf_multi_scf_get_input_charge": (C.c_int, [vp, vp]),
    "negf_tube_dos_complex": (C.c_int, [vp, C.c_int, vp, vp, vp, vp]),
    "negf_tube_dos_path": (C.c_int, [vp, C.c_int, C.c_int, vp, C.POINTER(C.c_int)]),
    "negf_write_transport_char": (C.c_int, [C.c_char_p, C.c_int, vp, vp, vp, vp, vp, vp]),
    "negf_write_ldos": (C.c_int, [C.c_char_p, C.c_int, vp, C.c_int, vp, vp]),
    "negf_batch_autotune_cached": (C.c_int, [vp, vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                             C.POINTER(C.c_int)]),
    "negf_rgf_diag_accumulate": (C.c_int, [C.POINTER(negf_chain_params), vp, C.c_int, vp, vp, vp, vp, vp]),
    "negf_rgf_diag_accumulate_noalloc": (C.c_int, [C.POINTER(negf_chain_params), vp, C.c_int, vp, vp, vp, vp,
                                                   vp]),
    "negf_rgf_noneq_accumulate": (C.c_int, [C.POINTER(negf_chain_params), vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp]),
    "negf_rgf_dos": (C.c_int, [C.POINTER(negf_chain_params), vp, C.c_int, vp, vp, vp, vp]),
    "negf_rgf_plan_create": (C.c_int, [C.POINTER(negf_chain_params), vp, C.c_int, vp, vp, C.c_int, C.POINTER(vp)]),
    "negf_rgf_plan_execute": (C.c_int, [vp, vp, vp, vp]),
    "negf_rgf_plan_status": (C.c_int, [vp, vp]),
    "negf_rgf_plan_enable_timing": (C.c_int, [vp, C.c_int]),
    "negf_rgf_plan_kernel_time": (C.c_int, [vp, dp, ip]),
    "negf_rgf_plan_destroy": (C.c_int, [vp]),
    "negf_gl_path": (C.c_int, [negf_complex, negf_complex, C.c_int, C.c_int, vp, dp, vp]),
    "negf_tube_params_default": (None, [C.POINTER(negf_tube_params)]),
    "negf_tube_create": (C.c_int, [C.POINTER(negf_tube_params), C.POINTER(vp)]),
    "negf_tube_destroy": (C.c_int, [vp]),
    "negf_tube_num_sites": (C.c_int, [vp]),
    "negf_tube_positions": (C.c_int, [vp, dp]),
    "negf_tube_modes": (C.c_int, [vp, ip, dp, ip]),
    "negf_tube_set_potential": (C.c_int, [vp, dp, dp]),
    "negf_tube_set_potential_device": (C.c_int, [vp, vp, vp]),
    "negf_tube_set_contacts": (C.c_int, [vp, dp, dp]),
    "negf_tube_status": (C.c_int, [vp]),
    "negf_tube_noneq_integrand": (C.c_int, [vp, C.c_int, dp, dp]),
    "negf_compute_rho0": (C.c_int, [vp, dp]),
    "negf_solve": (C.c_int, [vp, C.c_int, dp, ip]),
    "negf_solve_device": (C.c_int, [vp, C.c_int, vp, ip, vp]),
    "negf_get_charge_components": (C.c_int, [vp, dp, dp, dp]),
    "negf_tube_limits": (C.c_int, [vp, dp]),
    "negf_tube_noneq_layout": (C.c_int, [vp, ip, ip, dp]),
    "negf_tube_dos": (C.c_int, [vp, C.c_int, dp, dp, dp]),
    "negf_tube_current": (C.c_int, [vp, dp]),
    "negf_batch_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(vp)]),
    "negf_batch_destroy": (C.c_int, [vp]),
    "negf_batch_offsets": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
    "negf_batch_solve": (C.c_int, [vp, C.c_int, vp, vp]),
    "negf_batch_use_graph": (C.c_int, [vp, C.c_int]),
    "negf_batch_set_segment_length": (C.c_int, [vp, C.c_int]),
    "negf_batch_autotune": (C.c_int, [vp, vp, ip, dp]),
    "negf_batch_enable_timing": (C.c_int, [vp, C.c_int]),
    "negf_batch_last_time": (C.c_int, [vp, dp]),
    "negf_batch_energy_sites": (C.c_int, [vp, dp]),
    "negf_batch_status": (C.c_int, [vp]),
    "negf_batch_scf_create": (C.c_int, [vp, C.POINTER(vp), C.c_int, C.POINTER(negf_scf_params), C.POINTER(vp)]),
    "negf_batch_scf_destroy": (C.c_int, [vp]),
    "negf_batch_scf_set_external_potential": (C.c_int, [vp, dp]),
    "negf_batch_scf_set_input_charge": (C.c_int, [vp, dp]),
    "negf_batch_scf_get_input_charge": (C.c_int, [vp, dp]),
    "negf_batch_scf_use_graph": (C.c_int, [vp, C.c_int]),
    "negf_batch_scf_run": (C.c_int, [vp, dp, dp, dp, ip, dp, ip]),
    "negf_block_chain_create": (C.c_int, [C.c_int, C.c_int, dp, dp, C.POINTER(vp)]),
    "negf_block_chain_destroy": (C.c_int, [vp]),
    "negf_block_chain_set_potential": (C.c_int, [vp, dp, dp]),
    "negf_block_rgf": (C.c_int, [vp, C.c_int, vp, vp, C.c_double, C.c_int, dp, vp, ip]),
    "negf_block_self_energies": (C.c_int, [vp, C.c_int, vp, C.c_double, C.c_int, dp, ip]),
    "negf_block_probe": (C.c_int, [C.c_int, vp, vp, vp]),
    "negf_block_chain_set_scratch_limit": (C.c_int, [vp, C.c_size_t]),
    "negf_block_chain_dims": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "negf_block_material_create": (C.c_int, [vp, C.POINTER(negf_tube_params), C.c_double, C.c_double, C.c_int,
                                             C.POINTER(vp)]),
    "negf_block_material_destroy": (C.c_int, [vp]),
    "negf_block_material_set_potential": (C.c_int, [vp, dp, dp]),
    "negf_block_material_set_contacts": (C.c_int, [vp, dp, dp]),
    "negf_block_material_rho0": (C.c_int, [vp, dp]),
    "negf_block_solve": (C.c_int, [vp, dp, C.POINTER(C.c_int)]),
    "negf_block_scf_bias_point": (C.c_int, [vp, dp, dp, C.c_double, C.POINTER(C.c_int), C.POINTER(negf_scf_params),
                                            C.c_int, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "negf_block_last_time": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "negf_batch_scf_get_kernel": (C.c_int, [vp, dp]),
    "negf_batch_scf_set_preconditioner": (C.c_int, [vp, dp, dp]),
    "negf_batch_scf_set_surrogate": (C.c_int, [vp, vp, dp]),
    "negf_batch_scf_run_surrogate": (C.c_int, [vp, dp, C.c_int, C.c_int, dp, dp, ip, dp, ip, dp]),
    "negf_surrogate_create": (C.c_int, [C.c_int, C.c_ulonglong, C.POINTER(vp)]),
    "negf_surrogate_destroy": (C.c_int, [vp]),
    "negf_surrogate_num_params": (C.c_int, [vp]),
    "negf_surrogate_num_sites": (C.c_int, [vp]),
    "negf_surrogate_get_params": (C.c_int, [vp, C.POINTER(C.c_float), dp]),
    "negf_surrogate_set_params": (C.c_int, [vp, C.POINTER(C.c_float), dp]),
    "negf_surrogate_forward": (C.c_int, [vp, C.c_int, vp, vp, vp, vp]),
    "negf_surrogate_jvp": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp]),
    "negf_surrogate_gradient": (C.c_int, [vp, C.c_int, vp, vp, vp, C.POINTER(C.c_float), dp]),
    "negf_surrogate_fit_output_scale": (C.c_int, [vp, C.c_int, vp]),
    "negf_surrogate_train": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_ulonglong, dp]),
    "negf_broyden_create": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(vp)]),
    "negf_broyden_destroy": (C.c_int, [vp]),
    "negf_broyden_set_nin": (C.c_int, [vp, dp]),
    "negf_broyden_get_nin": (C.c_int, [vp, dp]),
    "negf_broyden_step": (C.c_int, [vp, dp, dp, dp, dp]),
    "negf_broyden_step_device": (C.c_int, [vp, vp, dp, dp, vp]),
    "negf_broyden_device_nin": (vp, [vp]),
    "negf_broyden_reset": (C.c_int, [vp, C.c_int, C.c_double]),
    "negf_broyden_get_step": (C.c_int, [vp]),
    "negf_scf_create": (C.c_int, [vp, C.POINTER(negf_scf_params), C.POINTER(vp)]),
    "negf_scf_destroy": (C.c_int, [vp]),
    "negf_scf_bias_point": (C.c_int, [vp, C.c_double, dp, dp, ip, dp]),
    "negf_scf_kernel": (C.c_int, [vp, dp]),
    "negf_scf_set_external_potential": (C.c_int, [vp, dp]),
    "negf_scf_use_graph": (C.c_int, [vp, C.c_int]),
    "negf_write_step_files": (C.c_int, [C.c_char_p, C.c_int, dp, dp, dp, dp]),
    "negf_write_qin": (C.c_int, [C.c_char_p, C.c_int, dp, dp]),
    "negf_write_charge_components": (C.c_int, [C.c_char_p, C.c_int, dp, dp, dp, dp, dp]),
    "negf_write_current_header": (C.c_int, [C.c_char_p]),
    "negf_write_current_line": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_double, dp,
                                          C.c_double, C.c_int, C.c_double]),
    "negf_read_charge_distribution": (C.c_int, [C.c_char_p, C.c_int, dp]),
    "negf_conductance": (C.c_int, [C.c_int, dp, dp, C.c_double, C.c_double, dp, dp]),
}

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m electrostatic_amd.build`"
            " (there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            raise RuntimeError(f"{LIB_PATH} does not export {name}: rebuild it")
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != NEGF_OK:
        msg = load().negf_last_error()
        raise NegfError(rc, msg.decode() if msg else "")
    return rc
