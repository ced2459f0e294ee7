"""The C-ABI library loads and exports every entry point include/gfx.h
declares; host-only entry points work without a device; device entry points
fail cleanly (error code, no crash) on bad arguments."""
import ctypes
import os
import re

import numpy as np
import pytest

from getfem_amd import _lib, core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gfx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared("gfx.h")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared("gfx.h")) == set(_lib.SIGNATURES)


def test_version_string():
    assert b"gfx950" in _lib.lib().gfx_version()


def test_null_arguments_rejected():
    L = _lib.lib()
    assert L.gfx_ctx_create(0, None) != 0
    assert L.gfx_plan_assemble(None) != 0
    assert L.gfx_mesh_fem_nb_dof(None) == -1
    n = ctypes.c_int()
    assert L.gfx_integration_rule(b"IM_NOPE(3)", ctypes.byref(n), None, None) == 2  # unsupported


def test_host_utilities_without_device():
    p, w = core.integration_rule("IM_TETRAHEDRON(5)")
    assert p.shape == (15, 3)
    v, g = core.fem_eval("FEM_PK(3,3)", [0.2, 0.2, 0.2])
    assert v.shape == (20,) and abs(v.sum() - 1) < 1e-13
    from getfem_amd import meshes
    pts, cvx = meshes.regular_unit_mesh(2)
    d, n = core.enumerate_dofs("GT_PK(3,1)", "FEM_PK(3,2)", 3, pts, cvx)
    assert n == 375 and d.shape == (48, 10)


def test_oracle_library_loads():
    import oracle
    assert oracle.lib().orc_im_npts(b"IM_TETRAHEDRON(8)") == 43
