"""ctypes binding of libgfx (the C ABI declared in include/gfx.h).

The library is built in-tree (``getfem_amd/_build/libgfx.so``).  There is no
fallback: if it is missing, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# GFX_LIBRARY names another build of the same library (same ABI)
LIB_PATH = os.environ.get("GFX_LIBRARY") or os.path.join(_HERE, "_build", "libgfx.so")

GFX_OK = 0
ERRORS = {
    1: "GFX_E_INVALID", 2: "GFX_E_UNSUPPORTED", 3: "GFX_E_DEGENERATE_ELEMENT",
    4: "GFX_E_OOM", 5: "GFX_E_HIP", 6: "GFX_E_STATE", 7: "GFX_E_PATTERN",
}
TERM_LAPLACE, TERM_ISO_ELAST, TERM_NEOHOOKEAN, TERM_ISOTROPIC = 0, 1, 2, 3
CSR, CSC = 0, 1
FORM_FACTORIZED, FORM_PERQP = 0, 1
PREC_FP64, PREC_FP32, PREC_BF16X2, PREC_FP64_MFMA, PREC_FP32_MFMA = 0, 1, 2, 3, 4
COEF_CONST, COEF_PER_ELEMENT = 0, 1
BUILD_MATRIX, BUILD_RHS, MODEL_LINEAR, MODEL_SYMMETRIC = 1, 2, 4, 8
PRECOND_BLOCK_JACOBI, PRECOND_ILDLT, PRECOND_ILU, PRECOND_DIAGONAL = 0, 1, 2, 3

# every symbol include/gfx.h declares, with (restype, argtypes)
_P = ctypes.c_void_p
_i32, _i64, _f32, _f64 = ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_double
_cs = ctypes.c_char_p
_pp = ctypes.POINTER(ctypes.c_void_p)


class GfxCoef(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("value", ctypes.c_double), ("h_values", ctypes.c_void_p)]


SIGNATURES = {
    "gfx_version": (_cs, []),
    "gfx_integration_rule": (_i32, [_cs, ctypes.POINTER(_i32), _P, _P]),
    "gfx_fem_eval": (_i32, [_cs, _P, ctypes.POINTER(_i32), _P, _P]),
    "gfx_enumerate_dofs": (_i32, [_cs, _cs, _i32, _i64, _P, _i64, _P, _P, ctypes.POINTER(_i64)]),
    "gfx_gwfl_compile": (_i32, [_cs, _cs, ctypes.c_char_p, _i64]),
    "gfx_gwfl_order": (_i32, [_cs, _cs, ctypes.POINTER(_i32)]),
    "gfx_gwfl_compile_scalar": (_i32, [_cs, _cs, _i32, _cs, ctypes.c_char_p, _i64]),
    "gfx_ctx_create": (_i32, [_i32, _pp]),
    "gfx_ctx_destroy": (_i32, [_P]),
    "gfx_ctx_set_stream": (_i32, [_P, _P]),
    "gfx_ctx_stream": (_P, [_P]),
    "gfx_ctx_synchronize": (_i32, [_P]),
    "gfx_ctx_last_error": (_cs, [_P]),
    "gfx_mesh_create": (_i32, [_P, _cs, _i64, _P, _i64, _P, _pp]),
    "gfx_mesh_set_points": (_i32, [_P, _P]),
    "gfx_mesh_destroy": (_i32, [_P]),
    "gfx_mesh_fem_create": (_i32, [_P, _cs, _i32, _pp]),
    "gfx_mesh_fem_create_from_dofs": (_i32, [_P, _cs, _i32, _P, _i64, _pp]),
    "gfx_mesh_fem_nb_dof": (_i64, [_P]),
    "gfx_mesh_fem_nb_dof_per_element": (_i32, [_P]),
    "gfx_mesh_fem_basic_dofs": (_i32, [_P, _P]),
    "gfx_mesh_fem_dof_nodes": (_i32, [_P, _P]),
    "gfx_mesh_fem_destroy": (_i32, [_P]),
    "gfx_mesh_im_create": (_i32, [_P, _cs, _pp]),
    "gfx_mesh_im_nb_points": (_i32, [_P]),
    "gfx_mesh_im_points": (_i32, [_P, _P, _P]),
    "gfx_mesh_im_destroy": (_i32, [_P]),
    "gfx_plan_create": (_i32, [_P, _P, _i32, ctypes.POINTER(GfxCoef), _i32, _pp]),
    "gfx_plan_set_coef": (_i32, [_P, _i32, ctypes.POINTER(GfxCoef)]),
    "gfx_plan_set_form": (_i32, [_P, _i32, _i32]),
    "gfx_plan_build_pattern": (_i32, [_P, ctypes.POINTER(_i64)]),
    "gfx_plan_assemble": (_i32, [_P]),
    "gfx_plan_set_state": (_i32, [_P, _P]),
    "gfx_plan_set_state_device": (_i32, [_P, _P]),
    "gfx_plan_pin_pattern": (_i32, [_P, _i32]),
    "gfx_plan_pattern_id": (_i32, [_P, ctypes.POINTER(_i64)]),
    "gfx_plan_spmv_add": (_i32, [_P, _P, _P]),
    "gfx_plan_add_to_matrix": (_i32, [_P, _i32, _i64, _P, _P, _P]),
    "gfx_plan_assemble_residual": (_i32, [_P]),
    "gfx_plan_residual": (_i32, [_P, _P]),
    "gfx_plan_spmv": (_i32, [_P, _P, _P]),
    "gfx_plan_dof_constraints": (_i32, [_P, _i64, _P, _P, _P, _P, _i32]),
    "gfx_plan_solve_cg": (_i32, [_P, _P, _P, _f64, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_f64)]),
    "gfx_plan_copy_residual_device": (_i32, [_P, _P]),
    "gfx_plan_dims": (_i32, [_P, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "gfx_plan_export_csr": (_i32, [_P, _P, _P, _P]),
    "gfx_plan_export_csc": (_i32, [_P, _P, _P, _P]),
    "gfx_plan_device_csr": (_i32, [_P, _pp, _pp, _pp]),
    "gfx_plan_copy_csr_device": (_i32, [_P, _P, _P, _P]),
    "gfx_plan_set_timing": (_i32, [_P, _i32]),
    "gfx_plan_last_timings": (_i32, [_P, ctypes.POINTER(_f32), ctypes.POINTER(_f32)]),
    "gfx_plan_timing_summary": (_i32, [_P, ctypes.POINTER(_i32), ctypes.POINTER(_f64),
                                       ctypes.POINTER(_f64)]),
    "gfx_plan_timing_calls": (_i32, [_P, _i32, ctypes.POINTER(_i32), _P, _P]),
    "gfx_plan_destroy": (_i32, [_P]),
    "gfx_ws_create": (_i32, [_P, _pp]),
    "gfx_ws_destroy": (_i32, [_P]),
    "gfx_ws_add_fem_variable": (_i32, [_P, _cs, _P, _i64, _P]),
    "gfx_ws_add_fem_constant": (_i32, [_P, _cs, _P, _P]),
    "gfx_ws_add_fixed_size_constant": (_i32, [_P, _cs, _P, _i64]),
    "gfx_ws_set_values": (_i32, [_P, _cs, _P]),
    "gfx_ws_set_values_device": (_i32, [_P, _cs, _P]),
    "gfx_ws_set_assembled_matrix": (_i32, [_P, _i32, _i64, _P, _P, _P]),
    "gfx_ws_device_csr": (_i32, [_P, ctypes.POINTER(_i64), ctypes.POINTER(_i64), _pp, _pp, _pp]),
    "gfx_ws_matrix_dims": (_i32, [_P, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "gfx_ws_export_csr": (_i32, [_P, _P, _P, _P]),
    "gfx_ws_export_csc": (_i32, [_P, _P, _P, _P]),
    "gfx_ws_nb_trees": (_i32, [_P]),
    "gfx_ws_add_expression": (_i32, [_P, _cs, _P, _i32]),
    "gfx_ws_set_form": (_i32, [_P, _i32, _i32]),
    "gfx_ws_compile": (_i32, [_P]),
    "gfx_ws_assembly": (_i32, [_P, _i32]),
    "gfx_ws_plan": (_i32, [_P, _pp]),
    "gfx_ws_assembled_vector": (_i32, [_P, _P]),
    "gfx_ws_vector_size": (_i32, [_P, ctypes.POINTER(_i64)]),
    "gfx_ws_assembled_potential": (_i32, [_P, ctypes.POINTER(_f64)]),
    "gfx_ws_copy_assembled_vector_device": (_i32, [_P, _P]),
    "gfx_ws_describe": (_i32, [_P, ctypes.c_char_p, _i64]),
    "gfx_ws_solve_cg": (_i32, [_P, _P, _P, _f64, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_f64)]),
    "gfx_ws_dof_constraints": (_i32, [_P, _i64, _P, _P, _P, _P, _i32]),
    "gfx_ws_ilu_factor": (_i32, [_P, ctypes.POINTER(_i64)]),
    "gfx_ws_ilu_apply": (_i32, [_P, _P, _P]),
    "gfx_ws_solve_gmres": (_i32, [_P, _i32, _P, _P, _f64, _i32, _i32, ctypes.POINTER(_i32),
                                  ctypes.POINTER(_f64)]),
    "gfx_compliance_create": (_i32, [_P, _P, _f64, _f64, _f64, _f64, _pp]),
    "gfx_compliance_eval": (_i32, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "gfx_compliance_eval_noalloc": (_i32, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "gfx_plan_solve_pcg": (_i32, [_P, _i32, _P, _P, _f64, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_f64)]),
    "gfx_plan_ildlt_factor": (_i32, [_P, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "gfx_plan_ildlt_values": (_i32, [_P, _P]),
    "gfx_plan_ildlt_apply": (_i32, [_P, _P, _P]),
    "gfx_compliance_status": (_i32, [_P]),
    "gfx_compliance_destroy": (_i32, [_P]),
}

_lib = None


class GfxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, code), msg))
        self.code = code


def _bind_hip_runtime():
    """libgfx and torch both need libamdhip64.so.7 and the process keeps the
    first one loaded.  torch only runs on the HIP runtime it ships, so load it
    first when it is installed: libgfx then shares torch's runtime, and torch
    streams are valid libgfx streams."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    global _lib
    if _lib is None:
        _bind_hip_runtime()
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libgfx is not built (%s missing): run __graft_entry__.build() or "
                "`make -C getfem_amd/csrc`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code, ctx=None):
    if code != GFX_OK:
        msg = lib().gfx_ctx_last_error(ctx).decode() if ctx else ""
        raise GfxError(code, msg)
