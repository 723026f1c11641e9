"""ctypes binding of the C ABI in include/efs.h (libefs_hip.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback for the step.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libefs_hip.so")

FIELD_VEL, FIELD_RHO, FIELD_GAMMA, FIELD_BETA, FIELD_DIVERGENCE, FIELD_PRESSURE, FIELD_FORCES, FIELD_BOUNDARY = range(8)
STAGE_VEL_DECAY, STAGE_VEL_EXTRAPOLATE, STAGE_RHO_SHARPEN, STAGE_RHO_ADVECT, STAGE_VEL_ADVECT, STAGE_ADD_FORCE, STAGE_PROJECT = range(7)
STAGE_NAMES = ["0 Vel Decay", "1 Vel Extrapolate", "2 Rho Sharpen", "3 Rho Advect", "4 Vel Advect", "5 Add Force", "6 Project"]
SOLVER_SELF_CG, SOLVER_JACOBI, SOLVER_MULTIGRID = 2, 3, 4
CRIT_REFERENCE, CRIT_RELATIVE = 0, 1
PRECOND_NONE, PRECOND_LEARNED, PRECOND_MG = 0, 1, 2
GUESS_DIVERGENCE, GUESS_PREVIOUS = 0, 1
KT_CG_DIR_SPMV, KT_CG_UPDATE, KT_JACOBI, KT_OTHER, KT_MG_DOWN, KT_ADV_BWD, KT_ADV_FWD, KT_WALLS_DIV, KT_GRADIENT = range(9)


class Params(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_float),
        ("gravity_force", ctypes.c_float),
        ("lambda_minus", ctypes.c_float),
        ("lambda_plus", ctypes.c_float),
        ("eta", ctypes.c_float),
        ("liquid_threshold", ctypes.c_float),
        ("rho_d", ctypes.c_float),
        ("vel_decay", ctypes.c_float),
        ("gravity_enabled", ctypes.c_int32),
        ("solver_type", ctypes.c_int32),
        ("max_steps", ctypes.c_int32),
        ("solver_threshold", ctypes.c_float),
        ("disable_mask", ctypes.c_uint32),
        ("criterion", ctypes.c_int32),
        ("precond", ctypes.c_int32),
        ("true_residual", ctypes.c_int32),
        ("initial_guess", ctypes.c_int32),
        ("re_calc_div", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("rr", ctypes.c_double),
        ("rel_residual", ctypes.c_double),
        ("true_rel_residual", ctypes.c_double),
        ("bnorm", ctypes.c_double),
        ("fallback", ctypes.c_int32),
        ("warm_started", ctypes.c_int32),
        ("stage_ms", ctypes.c_float * 7),
    ]


class LearnedOpts(ctypes.Structure):
    _fields_ = [
        ("out_scale", ctypes.c_float),
        ("jacobi_weight", ctypes.c_float),
        ("warm_start", ctypes.c_int32),
        ("max_iters", ctypes.c_int32),
        ("every", ctypes.c_int32),
        ("stall_window", ctypes.c_int32),
        ("stall_ratio", ctypes.c_float),
    ]


class QueryResult(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int32),
        ("max_pos", ctypes.c_float * 3),
        ("min_pos", ctypes.c_float * 3),
        ("avg_pos", ctypes.c_float * 3),
        ("vel", ctypes.c_float * 3),
        ("pressure", ctypes.c_float),
    ]


class Prim(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("solid", ctypes.c_int32),
        ("a", ctypes.c_double * 3),
        ("b", ctypes.c_double * 3),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "efs_default_params": (None, [ctypes.POINTER(Params)]),
    "efs_create_3d": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_void_p)]),
    "efs_create_2d": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_void_p)]),
    "efs_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "efs_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Params)]),
    "efs_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Params)]),
    "efs_available_solvers": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    "efs_dims": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "efs_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    "efs_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    "efs_upload_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    "efs_download_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    "efs_render_feed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "efs_render_feed_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "efs_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    "efs_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Stats)]),
    "efs_solve_pressure": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    "efs_project_divergence": (ctypes.c_int, [ctypes.c_void_p]),
    "efs_apply_operator": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "efs_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "efs_get_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "efs_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "efs_set_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "efs_set_kernel_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "efs_kernel_time": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]),
    "efs_scene_rasterise": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Prim), ctypes.c_int32]),
    "efs_scene_face_noise": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double]),
    "efs_scene_voxelise": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                                          ctypes.POINTER(ctypes.c_float), ctypes.c_int32, ctypes.c_float, ctypes.c_float]),
    "efs_scene_force_region": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int32,
                                              ctypes.POINTER(ctypes.c_int32)]),
    "efs_query_box": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                     ctypes.POINTER(QueryResult)]),
    "efs_query_box_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                     ctypes.POINTER(QueryResult)]),
    "efs_unet_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32), ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "efs_unet_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "efs_unet_num_params": (ctypes.c_int64, [ctypes.c_void_p]),
    "efs_unet_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "efs_unet_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "efs_unet_get_grads": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "efs_unet_forward_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]),
    "efs_unet_loss_grad_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    "efs_unet_adam_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    "efs_unet_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "efs_unet_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "efs_default_learned_opts": (None, [ctypes.POINTER(LearnedOpts)]),
    "efs_set_learned_preconditioner": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(LearnedOpts)]),
    "efs_learned_coarse_dims": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "efs_learned_coarse_pair": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.POINTER(ctypes.c_double)]),
    "efs_last_error": (ctypes.c_char_p, []),
    "efs_version": (ctypes.c_char_p, []),
}

_lib = None


def load(path: str = LIB_PATH):
    """Loads libefs_hip.so and declares every C-ABI signature; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"HIP backend library not found at {path}; build it with "
            "`python eulerianfluidsimulation_amd/build.py` (there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class EfsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"efs error {code}: {msg}")
        self.code = code


def check(rc, allowed=(0,)):
    if rc not in allowed:
        raise EfsError(rc, load().efs_last_error().decode(errors="replace"))
    return rc
