"""ctypes binding of libgtnufft.so (the C ABI declared in include/gt_nufft.h,
gt_cgsense.h, gt_radial.h, gt_buffer.h and gt_b1map.h).

The library is built in-tree by gadgetron_amd/csrc/Makefile.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgtnufft.so")

# Every entry point include/gt_nufft.h declares, in declaration order.
EXPORTED_SYMBOLS = [
    "gt_nufft_create",
    "gt_nufft_destroy",
    "gt_nufft_preprocess",
    "gt_nufft_compute",
    "gt_nufft_compute_noalloc",
    "gt_nufft_mult_MH_M",
    "gt_nufft_convolve",
    "gt_nufft_fft",
    "gt_nufft_deapodize",
    "gt_nufft_grad_traj",
    "gt_nufft_get_matrix_size",
    "gt_nufft_get_matrix_size_os",
    "gt_nufft_get_W",
    "gt_nufft_get_num_samples_frames",
    "gt_nufft_set_tiling",
    "gt_nufft_profile",
    "gt_nufft_profile_stages",
    "gt_nufft_profile_read",
    "gt_last_error",
    "gt_device_count",
    "gt_set_device",
    "gt_malloc",
    "gt_free",
    "gt_memcpy_htod",
    "gt_memcpy_dtoh",
    "gt_memcpy_dtod",
    "gt_memset",
    "gt_array_add",
    "gt_array_sqrt",
    "gt_array_scale",
    "gt_array_mul_weights",
    "gt_stream_synchronize",
    "gt_device_synchronize",
    # include/gt_cgsense.h
    "gt_sense_create",
    "gt_sense_destroy",
    "gt_sense_mult_M",
    "gt_sense_mult_MH",
    "gt_sense_mult_MH_M",
    "gt_cgsense_solve",
    "gt_cgsense_solve_image",
    "gt_cgsense_batch_create",
    "gt_cgsense_batch_run",
    "gt_cgsense_batch_destroy",
    "gt_cgsense_batch_branches",
    "gt_cgsense_batch_item_length",
    "gt_image_operator_compute",
    "gt_cgsense_precond_weights",
    # include/gt_radial.h
    "gt_radial_trajectory_golden_ratio_2d",
    "gt_radial_trajectory_fixed_angle_2d",
    "gt_radial_trajectory_variable_angle_2d",
    "gt_radial_dcw_golden_ratio_2d",
    "gt_radial_dcw_fixed_angle_2d",
    "gt_kooshball_trajectory_3d",
    "gt_sdc_estimate_dcw",
    # include/gt_buffer.h
    "gt_buffer_create",
    "gt_buffer_destroy",
    "gt_buffer_setup",
    "gt_buffer_set_dcw",
    "gt_buffer_clear",
    "gt_buffer_add_frame_data",
    "gt_buffer_get_accumulated_coil_images",
    "gt_buffer_get_normalization_factor",
    "gt_buffer_get_combined_coil_image",
    # include/gt_b1map.h
    "gt_estimate_b1_map",
]

# gt_status
GT_OK = 0
STATUS_NAMES = {
    0: "GT_OK",
    1: "GT_ERR_INVALID_ARGUMENT",
    2: "GT_ERR_NOT_PREPROCESSED",
    3: "GT_ERR_SHAPE",
    4: "GT_ERR_RANGE",
    5: "GT_ERR_UNSUPPORTED",
    6: "GT_ERR_DEVICE",
    7: "GT_ERR_FFT",
    8: "GT_ERR_OUT_OF_MEMORY",
    9: "GT_ERR_INTERNAL",
}


class GTError(RuntimeError):
    """Raised for a non-zero gt_status; mirrors the reference's std::runtime_error."""

    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class gt_nufft_desc(ctypes.Structure):
    _fields_ = [
        ("D", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("N", ctypes.c_size_t * 4),
        ("N_os", ctypes.c_size_t * 4),
        ("os_factor", ctypes.c_double),
        ("W", ctypes.c_double),
        ("conv_type", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


class gt_cg_reg(ctypes.Structure):
    _fields_ = [
        ("weight", ctypes.c_double),
        ("image", ctypes.c_void_p),
        ("image_len", ctypes.c_size_t),
    ]


class gt_cg_desc(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("tc_tolerance", ctypes.c_double),
        ("encoding_weight", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("precond", ctypes.c_void_p),
        ("precond_len", ctypes.c_size_t),
        ("regs", ctypes.POINTER(gt_cg_reg)),
        ("n_regs", ctypes.c_int),
        ("x0", ctypes.c_void_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing; build it with gadgetron_amd/csrc/Makefile "
            "(or __graft_entry__.build()).  There is no CPU fallback.")
    L = ctypes.CDLL(LIB_PATH)
    P, S, I, D = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
    U = ctypes.c_uint
    SP = ctypes.POINTER(ctypes.c_size_t)
    FP = ctypes.POINTER(ctypes.c_float)
    PP = ctypes.POINTER(ctypes.c_void_p)
    sigs = {
        "gt_nufft_create": [PP, ctypes.POINTER(gt_nufft_desc)],
        "gt_nufft_destroy": [P],
        "gt_nufft_preprocess": [P, P, S, S, I, P],
        "gt_nufft_compute": [P, P, SP, I, P, SP, I, P, S, I, P],
        "gt_nufft_compute_noalloc": [P, P, SP, I, P, SP, I, P, S, I, P],
        "gt_nufft_mult_MH_M": [P, P, P, SP, I, P, S, P],
        "gt_nufft_convolve": [P, P, SP, I, P, SP, I, I, I, P],
        "gt_nufft_fft": [P, P, SP, I, I, I, P],
        "gt_nufft_deapodize": [P, P, SP, I, I, P],
        "gt_nufft_grad_traj": [P, P, SP, I, P, SP, I, P, S, P, P],
        "gt_nufft_get_matrix_size": [P, SP],
        "gt_nufft_get_matrix_size_os": [P, SP],
        "gt_nufft_get_W": [P, ctypes.POINTER(D)],
        "gt_nufft_get_num_samples_frames": [P, SP, SP],
        "gt_nufft_set_tiling": [P, ctypes.POINTER(I), I],
        "gt_nufft_profile": [P, I],
        "gt_nufft_profile_stages": [P, U],
        "gt_nufft_profile_read": [P, ctypes.POINTER(D), ctypes.POINTER(I), I],
        "gt_last_error": [],
        "gt_device_count": [ctypes.POINTER(I)],
        "gt_set_device": [I],
        "gt_malloc": [PP, S],
        "gt_free": [P],
        "gt_memcpy_htod": [P, P, S, P],
        "gt_memcpy_dtoh": [P, P, S, P],
        "gt_memcpy_dtod": [P, P, S, P],
        "gt_memset": [P, I, S, P],
        "gt_array_add": [P, P, S, I, P],
        "gt_array_sqrt": [P, S, P],
        "gt_array_scale": [P, S, ctypes.c_float, P],
        "gt_array_mul_weights": [P, P, S, S, P],
        "gt_stream_synchronize": [P],
        "gt_device_synchronize": [],
        "gt_sense_create": [PP, P, P, S, P, S],
        "gt_sense_destroy": [P],
        "gt_sense_mult_M": [P, P, P, I, P],
        "gt_sense_mult_MH": [P, P, P, I, P],
        "gt_sense_mult_MH_M": [P, P, P, I, P],
        "gt_cgsense_solve": [P, P, P, ctypes.POINTER(gt_cg_desc), ctypes.POINTER(D), P],
        "gt_cgsense_solve_image": [P, P, P, ctypes.POINTER(gt_cg_desc), ctypes.POINTER(D), P],
        "gt_cgsense_batch_create": [PP, PP, PP, PP, I, ctypes.POINTER(gt_cg_desc), I, P],
        "gt_cgsense_batch_run": [P, P],
        "gt_cgsense_batch_destroy": [P],
        "gt_cgsense_batch_branches": [P, ctypes.POINTER(I)],
        "gt_cgsense_batch_item_length": [P, ctypes.POINTER(I)],
        "gt_image_operator_compute": [P, S, I, FP, P, P],
        "gt_cgsense_precond_weights": [P, S, S, P, D, P, P],
        "gt_radial_trajectory_golden_ratio_2d": [P, U, U, U, U, I, P],
        "gt_radial_trajectory_fixed_angle_2d": [P, U, U, U, ctypes.c_float, P],
        "gt_radial_trajectory_variable_angle_2d": [P, P, U, U, U, P],
        "gt_radial_dcw_golden_ratio_2d": [P, U, U, ctypes.c_float, ctypes.c_float, U, I, P],
        "gt_radial_dcw_fixed_angle_2d": [P, U, U, ctypes.c_float, ctypes.c_float, P],
        "gt_kooshball_trajectory_3d": [P, U, U, P, P],
        "gt_sdc_estimate_dcw": [P, S, I, SP, ctypes.c_float, U, ctypes.c_float, P, P, P],
        "gt_buffer_create": [PP],
        "gt_buffer_destroy": [P],
        "gt_buffer_setup": [P, I, SP, SP, ctypes.c_float, U, U, U],
        "gt_buffer_set_dcw": [P, P, S],
        "gt_buffer_clear": [P, P],
        "gt_buffer_add_frame_data": [P, P, S, P, ctypes.POINTER(I), P],
        "gt_buffer_get_accumulated_coil_images": [P, P, P],
        "gt_buffer_get_normalization_factor": [P, FP],
        "gt_buffer_get_combined_coil_image": [P, P, P, P],
        "gt_estimate_b1_map": [P, I, SP, U, I, P, P],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.gt_last_error.restype = ctypes.c_char_p
    _lib = L
    return _lib


def check(status):
    if status != GT_OK:
        msg = lib().gt_last_error()
        raise GTError(status, msg.decode() if msg else "")
    return status


def sizes(seq):
    seq = [int(v) for v in seq]
    return (ctypes.c_size_t * max(1, len(seq)))(*seq)
