"""getfem_amd -- MI355X-native engine for GetFEM's generic assembly hot path.

``ga_workspace::assembly`` of GetFEM (src/getfem_generic_assembly_workspace.cc:
791-936) on a subset of GWFL expressions, executed by hand-written HIP kernels
for gfx950 behind the C ABI of ``include/gfx.h`` (libgfx).
"""
from ._lib import (COEF_CONST, COEF_PER_ELEMENT, CSC, CSR, FORM_FACTORIZED, FORM_PERQP,  # noqa: F401
                   PREC_BF16X2, PREC_FP32, PREC_FP32_MFMA, PREC_FP64, PREC_FP64_MFMA, TERM_ISO_ELAST, TERM_ISOTROPIC, TERM_LAPLACE,
                   TERM_NEOHOOKEAN, PRECOND_BLOCK_JACOBI, PRECOND_ILDLT, PRECOND_ILU,
                   PRECOND_DIAGONAL, GfxError)
from .core import Context, Mesh, MeshFem, MeshIm, Plan  # noqa: F401
from . import meshes  # noqa: F401

__all__ = ["Context", "Mesh", "MeshFem", "MeshIm", "Plan", "meshes", "GfxError"]
