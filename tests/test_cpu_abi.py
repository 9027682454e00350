"""CPU tests of the C ABI and the host-side API mirror (no GPU compute)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import vfvm
from vfvm import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vfvm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vfvm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    names = _declared()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    # and the Python binding declares a signature for each
    assert {n for n, _, _ in L.SIGNATURES} == set(names)


_STRUCTS = [("vfvm_mesh_info", L.MeshInfo), ("vfvm_op", L.OpStruct), ("vfvm_solve_opts", L.SolveOpts),
            ("vfvm_solve_stats", L.SolveStats), ("vfvm_driver_stats", L.DriverStats),
            ("vfvm_adjoint_stats", L.AdjointStats), ("vfvm_adjoint_wrt", L.AdjointWrt),
            ("vfvm_adaptive_opts", L.AdaptiveOpts), ("vfvm_adaptive_stats", L.AdaptiveStats)]


def test_struct_layouts_match_header(tmp_path):
    """Sizes and field offsets of every ctypes structure against the C compiler's layout
    of include/vfvm.h (gcc on a generated probe)."""
    import subprocess
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "vfvm.h"', "int main(void) {"]
    for cname, cls in _STRUCTS:
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    for cname, cls in _STRUCTS:
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_version_and_error_string():
    assert b"gfx950" in L.lib.vfvm_version()
    with pytest.raises(vfvm.VFVChemError) as e:
        vfvm.Mesh.lattice(1)
    assert e.value.code == L.ERR_INVALID_ARGUMENT and "n must be" in str(e.value)


def test_null_arguments_rejected():
    assert L.lib.vfvm_mesh_lattice(5, 0, 0.3, None) == L.ERR_INVALID_ARGUMENT
    assert L.lib.vfvm_mesh_info_get(None, None) == L.ERR_INVALID_ARGUMENT
    assert L.lib.vfvm_solve(None, None, None, None) == L.ERR_INVALID_ARGUMENT
    assert L.lib.vfvm_assemble(None, None, 0, 0.0) == L.ERR_INVALID_ARGUMENT


def test_solve_opts_defaults_mirror_linear_gmres_new():
    o = vfvm.solve_opts()
    # LinearGmres::new: tol 1e-6, restart 50, max_outer 200, Jacobi (linear_gmres.rs:20-22,79)
    assert (o.tol, o.restart, o.max_outer, o.precond, o.precision) == (1e-6, 50, 200, vfvm.PRECOND_JACOBI, vfvm.FP64)


def test_linear_gmres_argument_errors():
    # linear_gmres.rs:26-37
    for args, code in [((0, 1e-6, 50, 200), L.ERR_INVALID_ARGUMENT), ((1, 0.0, 50, 200), L.ERR_LINEAR_NON_POSITIVE_TOLERANCE),
                       ((1, float("nan"), 50, 200), L.ERR_LINEAR_NON_POSITIVE_TOLERANCE),
                       ((1, 1e-6, 0, 200), L.ERR_LINEAR_NON_POSITIVE_RESTART),
                       ((1, 1e-6, 50, 0), L.ERR_LINEAR_NON_POSITIVE_MAX_OUTER)]:
        with pytest.raises(vfvm.VFVChemError) as e:
            vfvm.LinearGmres.new_with_options(*args)
        assert e.value.code == code
    assert vfvm.LinearGmres.new(1).opts["restart"] == 50


def test_operator_lists_follow_reference_order():
    """SteadyHeat::assemble_matrix order (steady_heat.rs:143-171): domain diffusion,
    production, source; Dirichlet; per flux boundary diffusion, production, source, Neumann."""
    class FakeVars:
        bnd = [0, 1, 2]
    ph = vfvm.SteadyHeat.new()
    ph.add_heat_dom(0, 0, 1, 2, 3)
    ph.add_temp_bnd(0, 0)
    ph.add_hflx_bnd(2, 1)
    ops = ph.operators(FakeVars())
    assert [o[0] for o in ops] == [vfvm.OP_SCL_DOM_DIFFUSION, vfvm.OP_SCL_DOM_PRODUCTION, vfvm.OP_SCL_DOM_SOURCE,
                                   vfvm.OP_SCL_BND_DIRICHLET, vfvm.OP_SCL_BND_DIFFUSION, vfvm.OP_SCL_BND_PRODUCTION,
                                   vfvm.OP_SCL_BND_SOURCE, vfvm.OP_SCL_BND_NEUMANN]
    th = vfvm.TransientHeat.new()
    th.add_heat_dom(0, 0, 4, 1, 2, 3)
    th.add_hflx_bnd(1, 1)
    ops = th.operators(FakeVars())
    assert ops[0][0] == vfvm.OP_SCL_DOM_TIME and ops[4][0] == vfvm.OP_SCL_BND_TIME


def test_device_context_needs_a_gpu_or_fails_loudly():
    """On a host without a GPU, creating a context returns VFVM_ERR_HIP (no silent CPU path)."""
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(vfvm.VFVChemError) as e:
        vfvm.Context(0)
    assert e.value.code == L.ERR_HIP
