"""ctypes binding of libmafx.so (include/mafx.h).  The library is the only compute path: a missing
library or a missing GPU raises instead of falling back to anything else."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libmafx.so")

MAFX_F64, MAFX_F32_FIELD, MAFX_F32_DEMAG, MAFX_F32_STENCILS = 0, 1, 2, 3
# precision names of the Python layer: "f32" is the FP32-field path (demag FFT and stencils in FP32);
# "f32_demag" keeps the stencils in FP64; "f32_stencils" keeps the demag in FP64
PRECISIONS = {"f64": MAFX_F64, "f32": MAFX_F32_FIELD, "f32_demag": MAFX_F32_DEMAG, "f32_stencils": MAFX_F32_STENCILS}
DEMAG, EXCHANGE, DMI, UNIAXIAL, ZEEMAN, RKKY = 1, 2, 3, 4, 5, 6
# mafx_term_desc.flags (include/mafx.h)
TERM_CACHING, TERM_VERBOSE, TERM_HOST_TENSOR, TERM_HOMOGENEOUS_CALLBACK, TERM_HOST_CALLBACK = 1, 2, 4, 8, 16
TERM_DEVICE_TENSOR, TERM_OCTANT_TENSOR, TERM_FULL_TENSOR = 32, 64, 128
ATOM_DIPOLE, ATOM_EXCHANGE, ATOM_DMI, ATOM_UNIAXIAL, ATOM_ZEEMAN = 11, 12, 13, 14, 15

# every entry point include/mafx.h declares
EXPORTS = [
    "mafx_last_error", "mafx_device_count", "mafx_demag_calc_N", "mafx_demag_cache_name", "mafx_ctx_tensor_source",
    "mafx_ctx_create", "mafx_ctx_destroy",
    "mafx_ctx_set_stream", "mafx_ctx_sync", "mafx_ctx_device_bytes", "mafx_ctx_set_time", "mafx_ctx_update_term",
    "mafx_ctx_set_material", "mafx_llg_set_alpha_field", "mafx_heff", "mafx_energy", "mafx_llg_rhs",
    "mafx_dot", "mafx_max_abs", "mafx_normalize", "mafx_mean_m", "mafx_llg_create", "mafx_llg_destroy",
    "mafx_llg_set_alpha", "mafx_llg_step", "mafx_llg_run_until", "mafx_llg_relax", "mafx_llg_integrate_dense", "mafx_llg_get_stats",
    "mafx_llg_set_controller_state", "mafx_llg_get_controller_state", "mafx_lbfgs_minimize", "mafx_prof_enable",
    "mafx_prof_reset", "mafx_prof_report", "mafx_net_create", "mafx_net_destroy", "mafx_net_param_count",
    "mafx_net_set_params", "mafx_net_get_params", "mafx_net_forward", "mafx_net_gradients", "mafx_net_train_step",
    "mafx_net_warm_start", "mafx_lbfgs_set_convergence_file", "mafx_lbfgs_set_finish_context", "mafx_mean_m_noalloc",
]


class Mesh(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


# int (*)(void* user, double t, const double* m, double* h_out, void* stream)
FIELD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p)


class TermDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("flags", C.c_int32), ("value", C.c_double),
                ("axis", C.c_double * 3), ("field", C.c_double * 3),
                ("array0", C.POINTER(C.c_double)), ("array1", C.POINTER(C.c_double)),
                ("index0", C.POINTER(C.c_uint32)), ("callback", FIELD_FN), ("callback_user", C.c_void_p)]


class LlgStats(C.Structure):
    _fields_ = [("h", C.c_double), ("h_last", C.c_double), ("err_last", C.c_double),
                ("accepted", C.c_int64), ("rejected", C.c_int64), ("heff_evals", C.c_int64),
                ("counter_hmin", C.c_int64), ("counter_hmax", C.c_int64),
                ("counter_maxscale", C.c_int64), ("counter_minscale", C.c_int64)]


class LbfgsStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("energy_evals", C.c_int64), ("final_energy", C.c_double),
                ("final_torque_rel", C.c_double), ("stop_reason", C.c_int32), ("finish_iterations", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libmafx.so is not built ({LIB_PATH}); run python magnum.af_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, dp, i64, u32 = C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.c_uint32
    L.mafx_last_error.restype = C.c_char_p
    L.mafx_device_count.argtypes = [C.POINTER(C.c_int)]
    L.mafx_demag_calc_N.argtypes = [C.POINTER(Mesh), dp, C.c_int]
    L.mafx_demag_cache_name.argtypes = [C.POINTER(Mesh), C.c_int, C.c_int, C.c_int, C.c_char_p, i64]
    L.mafx_ctx_tensor_source.argtypes = [vp, C.POINTER(C.c_int)]
    L.mafx_ctx_create.argtypes = [C.POINTER(Mesh), C.c_double, dp, C.POINTER(TermDesc), C.c_int, C.c_int,
                                  C.POINTER(vp)]
    L.mafx_ctx_destroy.argtypes = [vp]
    L.mafx_ctx_set_stream.argtypes = [vp, vp]
    L.mafx_ctx_sync.argtypes = [vp]
    L.mafx_ctx_device_bytes.argtypes = [vp, C.POINTER(i64)]
    L.mafx_ctx_set_time.argtypes = [vp, C.c_double]
    L.mafx_ctx_update_term.argtypes = [vp, C.c_int, C.POINTER(TermDesc)]
    L.mafx_ctx_set_material.argtypes = [vp, C.c_double, dp]
    L.mafx_llg_set_alpha_field.argtypes = [vp, dp]
    L.mafx_heff.argtypes = [vp, vp, u32, vp, C.c_int]
    L.mafx_energy.argtypes = [vp, vp, u32, dp, dp]
    L.mafx_llg_rhs.argtypes = [vp, C.c_double, C.c_double, vp, vp, vp]
    L.mafx_dot.argtypes = [vp, vp, vp, dp]
    L.mafx_max_abs.argtypes = [vp, vp, dp]
    L.mafx_normalize.argtypes = [vp, vp, C.c_int]
    L.mafx_mean_m.argtypes = [vp, vp, dp]
    L.mafx_llg_create.argtypes = [vp, C.c_double, C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.POINTER(vp)]
    L.mafx_llg_destroy.argtypes = [vp]
    L.mafx_llg_set_alpha.argtypes = [vp, C.c_double]
    L.mafx_llg_step.argtypes = [vp, vp, dp]
    L.mafx_llg_run_until.argtypes = [vp, vp, dp, C.c_double, i64, C.POINTER(i64)]
    L.mafx_llg_relax.argtypes = [vp, vp, dp, C.c_double, i64, i64, C.c_int, i64, C.POINTER(i64)]
    L.mafx_llg_integrate_dense.argtypes = [vp, vp, dp, C.c_double, C.c_double, dp, i64, C.POINTER(i64)]
    L.mafx_llg_get_stats.argtypes = [vp, C.POINTER(LlgStats)]
    L.mafx_llg_set_controller_state.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.mafx_llg_get_controller_state.argtypes = [vp, dp, dp, C.POINTER(C.c_int)]
    L.mafx_lbfgs_minimize.argtypes = [vp, vp, C.c_double, i64, C.c_int, C.c_double, dp, C.POINTER(LbfgsStats)]
    L.mafx_lbfgs_set_convergence_file.argtypes = [vp, C.c_char_p]
    L.mafx_lbfgs_set_finish_context.argtypes = [vp, vp]
    L.mafx_mean_m_noalloc.argtypes = [vp, i64, vp, vp]
    L.mafx_prof_enable.argtypes = [vp, C.c_int]
    L.mafx_prof_reset.argtypes = [vp]
    L.mafx_prof_report.argtypes = [vp, C.c_char_p, i64]
    fp = C.POINTER(C.c_float)
    L.mafx_net_create.argtypes = [vp, C.POINTER(vp)]
    L.mafx_net_destroy.argtypes = [vp]
    L.mafx_net_param_count.argtypes = [vp, C.POINTER(i64)]
    L.mafx_net_set_params.argtypes = [vp, fp]
    L.mafx_net_get_params.argtypes = [vp, fp]
    L.mafx_net_forward.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, vp]
    L.mafx_net_gradients.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, fp, dp]
    L.mafx_net_train_step.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, dp]
    L.mafx_net_warm_start.argtypes = [vp, C.POINTER(Mesh), vp, fp, vp]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().mafx_last_error()
        raise RuntimeError(f"mafx error {rc}: {msg.decode() if msg else ''}")
