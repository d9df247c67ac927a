"""ctypes mirror of include/clip_amd.h and loader of libclip_amd.so."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclip_amd.so")

MAX_DIM = 3
NUM_SIDES = 6
MAX_GEOMETRIES = 16
MAX_INLET_MODES = 8

# status codes
OK, ERR_ARG, ERR_PARSE, ERR_HIP, ERR_IO, ERR_STATE, ERR_DIVERGED, ERR_NOMEM = 0, -1, -2, -3, -4, -5, -6, -7

CASE = {"bubble": 0, "drop": 1, "rti": 2, "jet": 3}
SIDES = ["x-", "x+", "y-", "y+", "z-", "z+"]
BC = {"wall": 0, "slip wall": 1, "free convect": 2, "periodic": 3, "neumann": 4, "velocity": 5, "do nothing": 6}
GEOM = {"circle": 0, "sphere": 1, "square": 2, "cube": 3, "perturbation": 4}
PRECISION = {"f64": 0, "f32": 1, "mixed_f16": 2, "mixed_bf16": 3}
FIELD = {"phi": 0, "rho": 1, "p": 2, "u": 3, "mu": 4, "dphi": 5, "normal": 6,
         "f": 7, "g": 8, "f_post": 9, "g_post": 10}
LAYOUT = {"aos": 0, "soa": 1}
INTERIOR_X_FASTEST = 0  # interior arrays of the reverse-mode calls: i + nx (j + ny k)
INTERIOR_X_SLOWEST = 1  # (i ny + j) nz + k, numpy's C order for index [i, j, k]


class ClipConfig(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("case_type", ctypes.c_int),
        ("final_step", ctypes.c_uint),
        ("output_interval", ctypes.c_uint),
        ("report_interval", ctypes.c_uint),
        ("checkpoint_interval", ctypes.c_uint),
        ("checkpoint_copy", ctypes.c_uint),
        ("t_final", ctypes.c_double),
        ("N", ctypes.c_double * MAX_DIM),
        ("reference_length", ctypes.c_double),
        ("reference_velocity", ctypes.c_double),
        ("Re", ctypes.c_double),
        ("We", ctypes.c_double),
        ("Pe", ctypes.c_double),
        ("Ca", ctypes.c_double),
        ("Bo", ctypes.c_double),
        ("Mo", ctypes.c_double),
        ("rho_ratio", ctypes.c_double),
        ("mu_ratio", ctypes.c_double),
        ("gravity", ctypes.c_double * MAX_DIM),
        ("interface_width", ctypes.c_double),
        ("mobility", ctypes.c_double),
        ("amplitude", ctypes.c_double),
        ("has_sigma", ctypes.c_int),
        ("has_mu_h", ctypes.c_int),
        ("has_gravity_scale", ctypes.c_int),
        ("sigma_override", ctypes.c_double),
        ("mu_h_override", ctypes.c_double),
        ("gravity_scale", ctypes.c_double),
        ("rho_h", ctypes.c_double),
        ("rho_l", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("mu_h", ctypes.c_double),
        ("mu_l", ctypes.c_double),
        ("tau_h", ctypes.c_double),
        ("tau_l", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("bc_type", ctypes.c_int * NUM_SIDES),
        ("bc_present", ctypes.c_int * NUM_SIDES),
        ("bc_value", (ctypes.c_double * MAX_DIM) * NUM_SIDES),
        ("num_geometries", ctypes.c_int),
        ("geom_type", ctypes.c_int * MAX_GEOMETRIES),
        ("geom_id", ctypes.c_int * MAX_GEOMETRIES),
        ("geom_center", (ctypes.c_double * MAX_DIM) * MAX_GEOMETRIES),
        ("geom_length", (ctypes.c_double * MAX_DIM) * MAX_GEOMETRIES),
        ("geom_radius", ctypes.c_double * MAX_GEOMETRIES),
        ("geom_amplitude", ctypes.c_double * MAX_GEOMETRIES),
        ("inlet_modes", ctypes.c_int),
        ("inlet_mode_amp", ctypes.c_double * MAX_INLET_MODES),
        ("inlet_mode_phase", ctypes.c_double * MAX_INLET_MODES),
        ("has_mobility_override", ctypes.c_int),
        ("mobility_override", ctypes.c_double),
        ("slip_specular", ctypes.c_int),
    ]


class ClipInfo(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("q", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("ext", ctypes.c_uint * MAX_DIM),
        ("nodes", ctypes.c_ulonglong),
        ("interior", ctypes.c_ulonglong),
        ("device_bytes", ctypes.c_ulonglong),
        ("step", ctypes.c_ulonglong),
    ]


# Entry points declared in include/clip_amd.h: name -> (restype, argtypes)
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "clip_config_from_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ClipConfig)]),
    "clip_config_from_string": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ClipConfig)]),
    "clip_config_derive": (ctypes.c_int, [ctypes.POINTER(ClipConfig)]),
    "clip_create": (ctypes.c_int, [ctypes.POINTER(ClipConfig), ctypes.c_int, ctypes.POINTER(_P)]),
    "clip_set_stream": (ctypes.c_int, [_P, _P]),
    "clip_init_fresh": (ctypes.c_int, [_P]),
    "clip_step": (ctypes.c_int, [_P, ctypes.c_int]),
    "clip_init_from_phi": (ctypes.c_int, [_P, _D]),
    "clip_initial_phi": (ctypes.c_int, [ctypes.POINTER(ClipConfig), _D]),
    "clip_synchronize": (ctypes.c_int, [_P]),
    "clip_get_field": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D]),
    "clip_set_field": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D]),
    "clip_macroscopic": (ctypes.c_int, [_P]),
    "clip_geometry_sdf": (ctypes.c_int, [ctypes.POINTER(ClipConfig), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _D]),
    "clip_sum": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "clip_sum_noalloc": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint), _P, _P, _P]),
    "clip_check_health": (ctypes.c_int, [_P, _D]),
    "clip_checkpoint_save": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "clip_checkpoint_load": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "clip_checkpoint_rotate": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint]),
    "clip_get_info": (ctypes.c_int, [_P, ctypes.POINTER(ClipInfo)]),
    "clip_last_step_timing": (ctypes.c_int, [_P, _D, _D, _D, _D]),
    "clip_enable_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "clip_enable_graph": (ctypes.c_int, [_P, ctypes.c_int]),
    "clip_adjoint_forward": (ctypes.c_int, [_P, _D, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int]),
    "clip_adjoint_backward": (ctypes.c_int, [_P, _D, _D, _D, _D, ctypes.c_int]),
    "clip_adjoint_set_normal_floor": (ctypes.c_int, [_P, ctypes.c_double]),
    "clip_adjoint_set_nesting": (ctypes.c_int, [_P, ctypes.c_int]),
    "clip_adjoint_inlet_gradient": (ctypes.c_int, [_P, _D, ctypes.c_int]),
    "clip_write_vts": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "clip_periodic_boundary": (ctypes.c_int, [_P, ctypes.c_int, _P, _P]),
    "clip_wall_boundary": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "clip_slip_wall_boundary": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "clip_free_convect_boundary": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "clip_neumann_boundary": (ctypes.c_int, [_P, _P, _P]),
    "clip_velocity_boundary": (ctypes.c_int, [_P, _P, _P, _P]),
    "clip_mirror_boundary": (ctypes.c_int, [_P, _P]),
    "clip_copy_rate": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_int, _D]),
    "clip_destroy": (None, [_P]),
    "clip_last_error": (ctypes.c_char_p, []),
}

_lib = None


class ClipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[clip status {code}] {msg}")
        self.code = code


def lib_path():
    return LIB_PATH


def load(path=None):
    """Loads the HIP library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ClipError(ERR_STATE, f"{path} is missing: run `python -m clip_amd.build` first")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != OK:
        msg = load().clip_last_error()
        raise ClipError(rc, msg.decode() if msg else "")
    return rc
