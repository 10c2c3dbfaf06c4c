"""ctypes mirror of include/mpdpft.h (the binding a Python caller uses).

The library is loaded from mpdpft_amd/lib/libmpdpft.so (built in-tree by
mpdpft_amd/build.py).  A missing library raises: there is no CPU fallback.
"""
import ctypes
import os

MAX_SPECIES = 8
MAX_LIST = 16
MAX_MPP = 64
MAX_CENTRES = 256

OK = 0
EINVAL = -1
ENOMEM = -2
EHIP = -3
ENOTCONV = -4
ENOBRACKET = -5
EPULAY = -6
EIO = -7
ENODEV = -8

_d = ctypes.c_double
_i = ctypes.c_int


class Params(ctypes.Structure):
    _fields_ = [
        ("abi_version", _i),
        ("System", _i), ("Symmetry", _i), ("TaskType", _i), ("TaskParameterCount", _i),
        ("DIM", _i), ("Units", _i), ("edge", _d), ("steps", _i), ("S", _i),
        ("Abundances", _d * MAX_SPECIES), ("RelAcc", _d), ("InternalAcc", _d), ("K", _i),
        ("DensityExpression", _i), ("n_mpp", _i), ("mpp", _d * MAX_MPP), ("n_Mpp", _i),
        ("Mpp", _d * MAX_LIST), ("tauVec", _d * MAX_SPECIES), ("tVec", _d * MAX_SPECIES),
        ("TVec", _d * MAX_SPECIES), ("Environments", _i * MAX_SPECIES), ("stretchfactor", _d),
        ("Wall", _d), ("Noise", _d), ("InterpolVQ", _i), ("n_interactions", _i),
        ("Interactions", _i * MAX_LIST), ("incrementalV", _d), ("muVec", _d * MAX_SPECIES),
        ("DeltamuModifier", _d), ("Mixer", _d * 3), ("thetaVec", _d * MAX_SPECIES),
        ("SCcriterion", _d), ("maxSCcount", _i), ("method", _i), ("DelFieldMethod", _i),
        ("MovieQ", _i), ("n_schedule", _i), ("Schedule", _i * MAX_LIST), ("n_ekin", _i),
        ("EkinTypes", _i * MAX_LIST), ("alpha", _d), ("beta", _d), ("gamma", _d), ("gammaH", _d),
        ("degeneracy", _d), ("n_regularize", _i), ("regularize", _i * MAX_LIST),
        ("RegularizationThreshold", _d), ("Print", _i),
        ("TaperOff", _i), ("RegularizeCoulombSingularity", _i), ("SmearingExtentDx", _d),
        ("OutPrecision", _i), ("MaxIteration", _i),
        ("HartreeMode", _i), ("SmoothingKWidth", _d), ("ConvergenceMode", _i), ("Precision", _i),
        ("WriteFiles", _i), ("n_centres", _i), ("centres", (_d * 5) * MAX_CENTRES),
        ("n_hyperbox", _i), ("TaskHyperBox", _d * (2 * MAX_SPECIES)),
        ("NoiseSeed", ctypes.c_uint64), ("CubeFile", _i),
    ]


class AdjointInfo(ctypes.Structure):
    _fields_ = [("tol", _d), ("max_iterations", _i), ("iterations", _i), ("relative_residual", _d),
                ("dN_dmu", _d)]


class Result(ctypes.Structure):
    _fields_ = [
        ("S", _i), ("SCcount", _i), ("converged", _i), ("LinMix", _i), ("PulMix", _i),
        ("errorCount", _i), ("warningCount", _i), ("CC", _d), ("muVec", _d * MAX_SPECIES),
        ("FinalAbundances", _d * MAX_SPECIES), ("DeltamuModifier", _d), ("TotalEnergy", _d),
        ("DispersalEnergies", _d * MAX_SPECIES), ("EnvironmentalEnergies", _d * MAX_SPECIES),
        ("InteractionEnergies", _d * MAX_LIST), ("AuxEkin", (_d * 4) * MAX_SPECIES),
        ("solve_seconds", _d), ("bisection_steps", _i), ("fp32_iterations", _i),
        ("AuxEint", (_d * 4) * MAX_SPECIES), ("bracket_failures", _i), ("pulay_rejections", _i),
    ]


class StageStats(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_longlong), ("total_ms", _d), ("units", _d),
                ("bytes", _d), ("flops", _d)]


HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmpdpft.so")

# Every entry point declared in include/mpdpft.h.
EXPORTS = [
    "mpdpft_abi_version", "mpdpft_params_sizeof", "mpdpft_result_sizeof", "mpdpft_params_default",
    "mpdpft_parse_input", "mpdpft_create", "mpdpft_destroy", "mpdpft_last_error", "mpdpft_grid_size",
    "mpdpft_solve", "mpdpft_get_density", "mpdpft_get_potential", "mpdpft_get_environment",
    "mpdpft_set_potential", "mpdpft_fft", "mpdpft_delfield", "mpdpft_interaction_potential",
    "mpdpft_density", "mpdpft_run_dir", "mpdpft_enable_timing", "mpdpft_get_stage_stats",
    "mpdpft_import_v", "mpdpft_set_initial_potential", "mpdpft_set_break_file",
    "mpdpft_batch_create", "mpdpft_batch_solve", "mpdpft_batch_get_fields", "mpdpft_batch_seconds",
    "mpdpft_batch_phase_stats", "mpdpft_batch_kernel_stats", "mpdpft_batch_last_error", "mpdpft_batch_destroy",
    "mpdpft_adjoint", "mpdpft_centre_gradient", "mpdpft_centre_forces", "mpdpft_task_gradient", "mpdpft_set_centres",
    "mpdpft_loop_begin", "mpdpft_step", "mpdpft_energy", "mpdpft_enforce_constraints", "mpdpft_mix",
    "mpdpft_inject_fault", "mpdpft_batch_solve_noalloc", "mpdpft_set_data_dir",
]
BATCH_FP32 = 1
NOALLOC_SCRATCH_BYTES = 64 << 20
NOALLOC_MAX_INSTANCES = 4096
FAULT_DUPLICATE_PULAY = 1
FAULT_NAN_POTENTIAL = 2

_lib = None


def load_library(path=None):
    """Load libmpdpft.so (MPDPFT_LIBRARY names another build of it); raises
    OSError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("MPDPFT_LIBRARY") or LIB_PATH
    if not os.path.exists(path):
        raise OSError("libmpdpft.so not built (%s); run mpdpft_amd/build.py" % path)
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER(Params)
    R = ctypes.POINTER(Result)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    lib.mpdpft_abi_version.restype = _i
    lib.mpdpft_params_sizeof.restype = ctypes.c_size_t
    lib.mpdpft_result_sizeof.restype = ctypes.c_size_t
    lib.mpdpft_params_default.argtypes = [P]
    lib.mpdpft_params_default.restype = None
    lib.mpdpft_parse_input.argtypes = [ctypes.c_char_p, P]
    lib.mpdpft_parse_input.restype = _i
    lib.mpdpft_create.argtypes = [P, _i, ctypes.POINTER(vp)]
    lib.mpdpft_create.restype = _i
    lib.mpdpft_destroy.argtypes = [vp]
    lib.mpdpft_destroy.restype = None
    lib.mpdpft_last_error.argtypes = [vp]
    lib.mpdpft_last_error.restype = ctypes.c_char_p
    lib.mpdpft_grid_size.argtypes = [vp]
    lib.mpdpft_grid_size.restype = ctypes.c_size_t
    lib.mpdpft_solve.argtypes = [vp, R]
    lib.mpdpft_solve.restype = _i
    for n in ("mpdpft_get_density", "mpdpft_get_potential", "mpdpft_get_environment"):
        getattr(lib, n).argtypes = [vp, dp]
        getattr(lib, n).restype = _i
    lib.mpdpft_set_potential.argtypes = [vp, dp]
    lib.mpdpft_set_potential.restype = _i
    lib.mpdpft_fft.argtypes = [vp, _i, dp]
    lib.mpdpft_fft.restype = _i
    lib.mpdpft_delfield.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    lib.mpdpft_delfield.restype = _i
    lib.mpdpft_interaction_potential.argtypes = [vp, _i, _i, dp, dp]
    lib.mpdpft_interaction_potential.restype = _i
    lib.mpdpft_density.argtypes = [vp, _i, _d, dp, dp, dp]
    lib.mpdpft_density.restype = _i
    lib.mpdpft_enable_timing.argtypes = [vp, _i]
    lib.mpdpft_enable_timing.restype = _i
    lib.mpdpft_get_stage_stats.argtypes = [vp, ctypes.POINTER(StageStats), _i, ctypes.POINTER(_i)]
    lib.mpdpft_get_stage_stats.restype = _i
    lib.mpdpft_run_dir.argtypes = [ctypes.c_char_p, _i, R]
    lib.mpdpft_run_dir.restype = _i
    lib.mpdpft_import_v.argtypes = [ctypes.c_char_p, P, dp, ctypes.POINTER(_i)]
    lib.mpdpft_import_v.restype = _i
    lib.mpdpft_set_data_dir.argtypes = [ctypes.c_char_p]
    lib.mpdpft_set_data_dir.restype = _i
    lib.mpdpft_set_initial_potential.argtypes = [vp, dp]
    lib.mpdpft_set_initial_potential.restype = _i
    lib.mpdpft_set_break_file.argtypes = [vp, ctypes.c_char_p]
    lib.mpdpft_set_break_file.restype = _i
    lib.mpdpft_batch_create.argtypes = [P, _i, _i, _i, ctypes.POINTER(vp)]
    lib.mpdpft_batch_create.restype = _i
    lib.mpdpft_batch_solve.argtypes = [vp, R]
    lib.mpdpft_batch_solve.restype = _i
    lib.mpdpft_batch_get_fields.argtypes = [vp, _i, dp, dp]
    lib.mpdpft_batch_get_fields.restype = _i
    lib.mpdpft_batch_seconds.argtypes = [vp]
    lib.mpdpft_batch_seconds.restype = ctypes.c_double
    lib.mpdpft_batch_phase_stats.argtypes = [vp, vp, vp, vp]
    lib.mpdpft_batch_phase_stats.restype = ctypes.c_int
    lib.mpdpft_batch_kernel_stats.argtypes = [vp, vp, vp, vp, vp]
    lib.mpdpft_batch_kernel_stats.restype = ctypes.c_int
    lib.mpdpft_batch_last_error.argtypes = [vp]
    lib.mpdpft_batch_last_error.restype = ctypes.c_char_p
    lib.mpdpft_batch_destroy.argtypes = [vp]
    lib.mpdpft_batch_destroy.restype = None
    lib.mpdpft_batch_solve_noalloc.argtypes = [P, _i, _i, _i, vp, ctypes.c_size_t, R, ctypes.c_char_p,
                                               ctypes.c_size_t]
    lib.mpdpft_batch_solve_noalloc.restype = _i
    lib.mpdpft_adjoint.argtypes = [vp, dp, dp, ctypes.POINTER(AdjointInfo)]
    lib.mpdpft_adjoint.restype = _i
    lib.mpdpft_centre_gradient.argtypes = [vp, dp, dp]
    lib.mpdpft_centre_gradient.restype = _i
    lib.mpdpft_centre_forces.argtypes = [vp, dp]
    lib.mpdpft_centre_forces.restype = _i
    lib.mpdpft_task_gradient.argtypes = [vp, dp]
    lib.mpdpft_task_gradient.restype = _i
    lib.mpdpft_set_centres.argtypes = [vp, _i, dp]
    lib.mpdpft_set_centres.restype = _i
    lib.mpdpft_loop_begin.argtypes = [vp]
    lib.mpdpft_loop_begin.restype = _i
    lib.mpdpft_step.argtypes = [vp, R, ctypes.POINTER(_i)]
    lib.mpdpft_step.restype = _i
    lib.mpdpft_energy.argtypes = [vp, R]
    lib.mpdpft_energy.restype = _i
    lib.mpdpft_enforce_constraints.argtypes = [vp, dp, dp]
    lib.mpdpft_enforce_constraints.restype = _i
    lib.mpdpft_mix.argtypes = [vp, _i, dp]
    lib.mpdpft_mix.restype = _i
    lib.mpdpft_inject_fault.argtypes = [vp, _i, _i]
    lib.mpdpft_inject_fault.restype = _i
    if lib.mpdpft_params_sizeof() != ctypes.sizeof(Params):
        raise OSError("mpdpft_params layout mismatch: C %d vs ctypes %d"
                      % (lib.mpdpft_params_sizeof(), ctypes.sizeof(Params)))
    if lib.mpdpft_result_sizeof() != ctypes.sizeof(Result):
        raise OSError("mpdpft_result layout mismatch")
    _lib = lib
    return lib
