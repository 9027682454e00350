"""vfvm_solve_device_noalloc (the solve of the assembled system with no device allocation)
against vfvm_solve_device on the same systems: two calls in a row on one GPU, the second on a
smaller system, each within 1e-12 (relative) of the allocating entry point's solution; sizes
beyond the fixed scratch array and unsupported options are rejected."""
import numpy as np
import pytest
import torch

import vfvm
from problems import heat_c1, mass_c2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


def _pair(ctx, p, **opts):
    p.to_device(ctx)
    ctx.pattern_build(p.ops)
    ctx.assemble(p.ops)
    x_ref = torch.zeros(p.N, dtype=torch.float64, device="cuda")
    x_na = torch.zeros(p.N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st_ref = ctx.solve_device(x_ref.data_ptr(), **opts)
    st_na = ctx.solve_device_noalloc(x_na.data_ptr(), **opts)
    torch.cuda.synchronize()
    a, b = x_ref.cpu().numpy(), x_na.cpu().numpy()
    return a, b, st_ref, st_na


@pytest.mark.parametrize("precond", [vfvm.PRECOND_JACOBI, vfvm.PRECOND_NONE])
def test_noalloc_matches_solve_device_twice(ctx, precond):
    opts = dict(tol=1e-11, restart=40, max_outer=400, precision=vfvm.FP64, precond=precond)
    for n, make in ((14, mass_c2), (7, heat_c1)):  # the second call on a smaller system
        print(f"[no-alloc solve n={n} {make.__name__} precond={precond}]", flush=True)
        a, b, st_ref, st_na = _pair(ctx, make(vfvm.Mesh.lattice(n)), **opts)
        assert st_na.rel_residual <= 1e-11 and st_ref.rel_residual <= 1e-11
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a)), np.max(np.abs(a - b))
        assert st_na.iterations == st_ref.iterations
        print("  bitwise equal:", bool(np.array_equal(a, b)), "steps", st_na.iterations)


def test_noalloc_rejects_what_it_cannot_hold(ctx):
    p = heat_c1(vfvm.Mesh.lattice(40))  # 64000 unknowns x 257 basis vectors > 16 Mi doubles
    p.to_device(ctx)
    ctx.pattern_build(p.ops)
    ctx.assemble(p.ops)
    x = torch.zeros(p.N, dtype=torch.float64, device="cuda")
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.solve_device_noalloc(x.data_ptr(), tol=1e-8, restart=256, precond=vfvm.PRECOND_JACOBI)
    assert e.value.code == 1 and "workspace" in str(e.value)
    for bad in (dict(precision=vfvm.FP32_REFINED), dict(precond=vfvm.PRECOND_ILU0)):
        with pytest.raises(vfvm.VFVChemError) as e:
            ctx.solve_device_noalloc(x.data_ptr(), tol=1e-8, restart=20, **bad)
        assert e.value.code == 1
    # the context still solves after the rejections
    st = ctx.solve_device_noalloc(x.data_ptr(), tol=1e-8, restart=30, max_outer=400, precond=vfvm.PRECOND_JACOBI)
    assert st.rel_residual <= 1e-8
