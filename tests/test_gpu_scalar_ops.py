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


def test_transient_hook_weight_enters_time_previous(ctx):
    """TransientBase::solve updates the previous values of every field after each step
    (transient_base.rs:131, _update_prev :297-311), so a weight that the update hook computes
    from the unknown (rho c_p = 1 + T / 2 here) reaches the next step's OpSclDomTimePrevious
    terms with its value at the end of the step.  Oracle: the same implicit-Euler steps with
    SuperLU, the weight's previous values taken from the previous state."""
    print("[transient hook-driven time-previous weight]", flush=True)
    mesh = vfvm.Mesh.lattice(7, 1, 0.3)
    p = heat_flow_c5like(mesh, transient=True)
    p.to_device(ctx)
    rcp = p.rcp_id
    T_id = p.unknown

    def weight(T):
        return 1.0 + 0.5 * T

    def hook(t):
        ctx.scldom_set(rcp, weight(ctx.scldom_get(T_id)))

    dt, steps = 0.02, 3
    ctx.set_update_hook(hook)
    try:
        ctx.transient_solve(p.ops, dt, steps, 20, 1e-12, 1.0, lin=dict(tol=1e-13, restart=100, max_outer=400,
                                                                        precond=vfvm.PRECOND_ILU0))
    finally:
        ctx.set_update_hook(None)
    T_dev = ctx.scldom_get(T_id)

    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    T = np.zeros(p.N)
    for b in (0, 1):  # Dirichlet values at t = 0 (x = 0: 1, x = 1: 0)
        T[mesh.num_pts + mesh.bnd_nodes(b)] = 1.0 if b == 0 else 0.0
    T_prev, w_prev = T.copy(), weight(T)
    kind = p.scl[rcp][0]
    for _ in range(steps):
        # Picard (damping 1): the hook sets the weight from the iterate before every assembly;
        # the time-previous terms read the weight's and the unknown's previous values
        Ti = T_prev.copy()
        for _it in range(50):
            p.scl[rcp] = (kind, weight(Ti))
            rp, col, val, b, _ = p.oracle_system(g, value_prev={T_id: T_prev, rcp: w_prev}, dt=dt, unknown_value=Ti)
            A = sp.csr_matrix((val, col, rp), shape=(p.N, p.N)).tocsc()
            Tn = spla.spsolve(A, b)
            done = np.max(np.abs(Tn - Ti)) <= 1e-14 * np.max(np.abs(Tn))
            Ti = Tn
            if done:
                break
        T_prev, w_prev = Ti, weight(Ti)
    assert np.max(np.abs(T_dev - T_prev)) <= 1e-9 * np.max(np.abs(T_prev))
