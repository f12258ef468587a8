"""ctypes binding of libbeta1_mhd.so (declarations in include/beta1_mhd.h)."""
import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbeta1_mhd.so")


class MhdError(RuntimeError):
    """A C-ABI call returned a negative status (``code``; MHD_ERR_DT = -3)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class MhdOptions(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("recon", C.c_int), ("ghost", C.c_int),
                ("hbm_budget_gb", C.c_double), ("autotune", C.c_int), ("xcd_order", C.c_int), ("overlap", C.c_int)]


class MhdOTDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("gamma", C.c_double), ("cfl", C.c_double), ("seed", C.c_ulonglong),
                ("noise", C.c_double)]


class MhdWindDesc(C.Structure):
    _fields_ = [("nang", C.c_int), ("nrad", C.c_int), ("r_in", C.c_double), ("r_out", C.c_double),
                ("gamma", C.c_double), ("cfl", C.c_double), ("n_cc", C.c_double), ("v_kms", C.c_double),
                ("vdir", C.c_double * 3), ("T_eV", C.c_double), ("B_nT", C.c_double * 3), ("m_amu", C.c_double)]


class MhdDipoleDesc(C.Structure):
    _fields_ = [("B_eq_nT", C.c_double), ("offset_z", C.c_double), ("softening", C.c_double),
                ("single_fluid", C.c_int), ("no_ambipolar", C.c_int), ("no_hall", C.c_int),
                ("hall_r_min", C.c_double), ("hall_r_max", C.c_double), ("local_eta_max", C.c_double),
                ("local_center", C.c_double * 3), ("local_r_decay", C.c_double), ("local_r_cutoff", C.c_double),
                ("no_resistance", C.c_int), ("resist_subcycles", C.c_int), ("resist_eta_max", C.c_double)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PI = C.POINTER(C.c_int)
_PD = C.POINTER(C.c_double)
_S = C.c_size_t

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "mhd_create": (_I, [C.POINTER(MhdOptions), C.POINTER(_P)]),
    "mhd_destroy": (_I, [_P]),
    "mhd_last_error": (C.c_char_p, [_P]),
    "mhd_add_case_orszag_tang": (_I, [_P, C.POINTER(MhdOTDesc), _PI]),
    "mhd_add_case_lunar": (_I, [_P, C.POINTER(MhdWindDesc), _PI]),
    "mhd_add_case_mercury": (_I, [_P, C.POINTER(MhdWindDesc), C.POINTER(MhdDipoleDesc), _PI]),
    "mhd_adopt_host_case": (_I, [_P, _P, _PI]),
    "mhd_add_case_from_dir": (_I, [_P, C.c_char_p, C.POINTER(MhdWindDesc), _PI]),
    "mhd_refresh": (_I, [_P]),
    "mhd_write_restart": (_I, [_P, _I, C.c_char_p, C.c_longlong, _D]),
    "mhd_write_tecplot": (_I, [_P, _I, C.c_char_p, _D]),
    "mhd_read_restart": (_I, [_P, _I, C.c_char_p, C.POINTER(C.c_longlong), _PD]),
    "mhd_host_case_write_dir": (_I, [_P, C.c_char_p]),
    "mhd_host_case_from_dir": (_I, [C.c_char_p, C.POINTER(MhdWindDesc), _I, C.POINTER(_P)]),
    "mhd_host_write_restart": (_I, [_P, C.c_char_p, C.c_longlong, _D]),
    "mhd_host_write_tecplot": (_I, [_P, C.c_char_p, _D]),
    "mhd_host_read_restart": (_I, [_P, C.c_char_p, C.POINTER(C.c_longlong), _PD]),
    "mhd_finalize": (_I, [_P]),
    "mhd_step": (_I, [_P, _I]),
    "mhd_step_graph": (_I, [_P, _I]),
    "mhd_compute_dt": (_I, [_P, _I, _PD]),
    "mhd_last_dt": (_I, [_P, _I, _PD]),
    "mhd_divb_max": (_I, [_P, _I, _PD]),
    "mhd_synchronize": (_I, [_P]),
    "mhd_case_info": (_I, [_P, _I, _PI, _PI, _PI, _PI, _PI]),
    "mhd_block_info": (_I, [_P, _I, _I, _PI, _PI, _PI]),
    "mhd_host_block_info": (_I, [_P, _I, _PI, _PI, _PI]),
    "mhd_field_count": (_I, [_P, _I, _I, C.POINTER(_S)]),
    "mhd_download": (_I, [_P, _I, _I, _I, _PD, _S]),
    "mhd_upload": (_I, [_P, _I, _I, _I, _PD, _S]),
    "mhd_get_faces": (_I, [_P, _I, _I, _PI]),
    "mhd_get_nodes": (_I, [_P, _I, _I, _I, _PD, _S]),
    "mhd_get_constants": (_I, [_P, _I, _PD]),
    "mhd_profile": (_I, [_P, _I]),
    "mhd_profile_read": (_I, [_P, _I, _PD, _PI, C.POINTER(C.c_char_p)]),
    "mhd_memory": (_I, [_P, _PD]),
    "mhd_retire": (_I, [_P, _I]),
    "mhd_case_steps": (_I, [_P, _I, C.POINTER(C.c_longlong)]),
    "mhd_case_error": (_I, [_P, _I, _PI]),
    "mhd_tuning": (_I, [_P, _I, _PI, _PI, _PI]),
    "mhd_set_tuning": (_I, [_P, _I, _I, _I, _I]),
    "mhd_autotune": (_I, [_P]),
    "mhd_diagnostics": (_I, [_P, _I, _PD]),
    "mhd_diagnostics_noalloc": (_I, [_P, _PI, _I, _PD]),
    "mhd_host_case_ot": (_I, [C.POINTER(MhdOTDesc), _I, C.POINTER(_P)]),
    "mhd_host_case_lunar": (_I, [C.POINTER(MhdWindDesc), _I, C.POINTER(_P)]),
    "mhd_host_case_mercury": (_I, [C.POINTER(MhdWindDesc), C.POINTER(MhdDipoleDesc), _I, C.POINTER(_P)]),
    "mhd_host_case_destroy": (_I, [_P]),
    "mhd_host_case_info": (_I, [_P, _PI, _PI, _PI, _PI, _PI]),
    "mhd_host_field_count": (_I, [_P, _I, C.POINTER(_S)]),
    "mhd_host_download": (_I, [_P, _I, _I, _PD, _S]),
    "mhd_host_upload": (_I, [_P, _I, _I, _PD, _S]),
    "mhd_host_faces": (_I, [_P, _I, _PI]),
    "mhd_host_face_parts": (_I, [_P, _I, _I, _PI, _I, _PI]),
    "mhd_host_nodes": (_I, [_P, _I, _I, _PD, _S]),
    "mhd_host_constants": (_I, [_P, _PD]),
    "mhd_host_constants_two_fluid": (_I, [_P, _PD]),
    "mhd_host_apply_sync": (_I, [_P, _I]),
    "mhd_host_op_count": (_I, [_P, _I, C.POINTER(_S)]),
    "mhd_host_ops_independent": (_I, [_P, _I, C.POINTER(_I)]),
    "mhd_host_emf_shell_check": (_I, [_P, C.POINTER(_I)]),
    "mhd_case_begin": (_I, [_P, _I, _PD, _PI]),
    "mhd_add_block": (_I, [_P, _I, _I, _I, _I, _I, _PD, _PD, _PD, _S, C.POINTER(_PD)]),
    "mhd_set_topology": (_I, [_P, _I, _I, _PI]),
    "mhd_set_face_parts": (_I, [_P, _I, _I, _I, _I, _PI]),
    "mhd_case_end": (_I, [_P, _I]),
    "mhd_host_case_begin": (_I, [_I, _I, _PD, C.POINTER(_P)]),
    "mhd_host_add_block": (_I, [_P, _I, _I, _I, _I, _PD, _PD, _PD, _S, C.POINTER(_PD)]),
    "mhd_host_set_topology": (_I, [_P, _I, _PI]),
    "mhd_host_set_face_parts": (_I, [_P, _I, _I, _I, _PI]),
    "mhd_host_case_end": (_I, [_P]),
    "mhd_host_last_error": (C.c_char_p, [_P]),
    "mhd_step_dt": (_I, [_P]),
    "mhd_fluid_rhs": (_I, [_P]),
    "mhd_face_emf": (_I, [_P]),
    "mhd_edge_emf": (_I, [_P]),
    "mhd_curl_update": (_I, [_P]),
    "mhd_apply_update": (_I, [_P]),
    "mhd_sync": (_I, [_P, C.c_char_p]),
    "mhd_derived": (_I, [_P]),
}

_lib = None


def lib():
    """Load libbeta1_mhd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MhdError(f"{LIB_PATH} is missing: run beta1_amd.build.build() (no fallback path exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
