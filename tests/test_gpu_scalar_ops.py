"""GPU parity of OpSclDomAdvection / OpSclBndAdvection, OpSclDomTimePrevious and
OpSclBndTimeUnity (heat carried by a prescribed flow, steady and one implicit-Euler step):
CSR pattern, values and b bitwise equal to the oracle's; the transient solve within 1e-10 of
SuperLU."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import vfvm
import vfvm_oracle as O
from problems import heat_flow_c5like

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("n,transient", [(6, False), (9, False), (7, True)])
def test_heat_flow_assembly_vs_oracle(ctx, n, transient):
    print(f"[heat flow assemble n={n} transient={transient}]", flush=True)
    mesh = vfvm.Mesh.lattice(n, 3, 0.3)
    p = heat_flow_c5like(mesh, transient)
    p.to_device(ctx)
    dt = 0.02 if transient else 0.0
    ctx.pattern_build(p.ops)
    ctx.assemble(p.ops, dt)
    rp, col = ctx.pattern_get()
    vals, b = ctx.matrix_get()
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    scl, _, _, _, _ = p.oracle_inputs(g)
    prev = {p.unknown: np.zeros(p.N), p.rcp_id: scl[p.rcp_id][0]} if transient else None
    orp, ocol, oval, ob, _ = p.oracle_system(g, value_prev=prev, dt=dt)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(vals, oval), int(np.count_nonzero(vals != oval))
    assert np.array_equal(b, ob), int(np.count_nonzero(b != ob))
    x, st = ctx.solve(tol=1e-12, restart=50, max_outer=400, precond=vfvm.PRECOND_ILU0)
    A = sp.csr_matrix((oval, ocol, orp), shape=(p.N, p.N)).tocsc()
    xl = spla.spsolve(A, ob)
    assert np.linalg.norm(x - xl) <= 1e-10 * np.linalg.norm(xl)
