"""Parity of the interface operators (opscl_itf_*.rs) on joined multi-domain meshes, on the GPU.

Bars: CSR pattern identical to the oracle's per-domain assembly (assemble_domains, which
restates the reference's per-domain unknowns and its interface operators); values and b equal to
1e-14 relative (expected bitwise: the interface kernel adds every push in the reference's
order); solves within 1e-10 of SuperLU on the oracle's matrix, and the piecewise-linear
conduction profile through slabs of different conductivity to 1e-10.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import vfvm
from problems import X0, X1, itf_advdiff, itf_heat, itf_transient

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


def _check_system(ctx, p, dt=0.0, value_prev=None, unknown_value=None):
    ctx.pattern_build(p.ops)
    ctx.assemble(p.ops, dt)
    rp, col = ctx.pattern_get()
    vals, b = ctx.matrix_get()
    orp, ocol, oval, ob, _ = p.oracle_system(dt=dt, value_prev=value_prev, unknown_value=unknown_value)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.abs(vals - oval).max() <= 1e-14 * np.abs(oval).max()
    assert np.abs(b - ob).max() <= 1e-14 * max(np.abs(ob).max(), 1.0)
    print(" bitwise-equal values:", np.array_equal(vals, oval), "b:", np.array_equal(b, ob), flush=True)
    return orp, ocol, oval, ob


def _field_fns():
    k = lambda x: 1.0 + 0.5 * np.sin(2 * x[:, 0]) + 0.3 * x[:, 1] * x[:, 2]
    prd = lambda x: -0.2 - 0.1 * x[:, 0] * x[:, 1]
    q = lambda x: 1.0 + np.cos(3 * x[:, 2]) + x[:, 0]
    return k, prd, q


@pytest.mark.parametrize("n,splits", [(7, [3]), (9, [2, 6]), (12, [5])])
def test_itf_heat_pattern_values(ctx, n, splits):
    print(f"[interface heat assemble n={n} splits={splits}]", flush=True)
    k, prd, q = _field_fns()
    p = itf_heat(n, splits=splits, k=k, prd=prd, q=q)
    p.to_device(ctx)
    _check_system(ctx, p)


@pytest.mark.parametrize("previous", [False, True])
def test_itf_transient_pattern_values(ctx, previous):
    print(f"[interface transient assemble previous={previous}]", flush=True)
    p = itf_transient(8, m=(1.0, 2.5), s=(-0.5, -2.0), dt=0.1, previous=previous)
    p.to_device(ctx)
    # previous values: the unknown's (and the weight's) current values copied, then new values
    rng = np.random.default_rng(3)
    u_prev = 1.0 + 0.1 * rng.standard_normal(p.N)
    ctx.scldom_set(p.unknown, u_prev)
    ctx.scldom_update_prev(p.unknown)
    prev = {p.unknown: u_prev}
    if previous:
        w_prev = 2.0 + 0.3 * rng.standard_normal(p.N)
        ctx.scldom_set(1, w_prev)
        ctx.scldom_update_prev(1)
        prev[1] = w_prev
    # the unknown holds u_prev too (the production's lagged term reads its gradient)
    _check_system(ctx, p, dt=p.dt, value_prev=prev, unknown_value=u_prev)


@pytest.mark.parametrize("weighted", [False, True])
def test_itf_advection_pattern_values(ctx, weighted):
    print(f"[interface advection assemble weighted={weighted}]", flush=True)
    p = itf_advdiff(9, 5.0, splits=[3, 5], weighted=weighted)
    p.to_device(ctx)
    _check_system(ctx, p)


@pytest.mark.parametrize("precond", [vfvm.PRECOND_JACOBI, vfvm.PRECOND_ILU0])
def test_itf_piecewise_linear_solve(ctx, precond):
    print(f"[interface conduction solve pc={precond}]", flush=True)
    ks = (2.0, 0.5, 5.0)
    p = itf_heat(13, splits=[4, 8], k=ks)
    p.to_device(ctx)
    orp, ocol, oval, ob = _check_system(ctx, p)
    x, st = ctx.solve(None, tol=1e-13, restart=50, max_outer=400, precond=precond)
    A = sp.csr_matrix((oval, ocol, orp), shape=(p.N, p.N)).tocsc()
    x_lu = spla.spsolve(A, ob)
    assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) <= 1e-10
    # closed form: linear in each slab, continuous temperature and flux
    xs = p.pos[:, 0]
    cuts = [q.get("pos")[:, 0].min() for q in p.parts] + [1.0]
    res = np.array([(cuts[d + 1] - cuts[d]) / ks[d] for d in range(3)])
    flux = 1.0 / res.sum()
    exact = np.empty(p.N)
    t0 = 0.0
    for d in range(3):
        s = p.dom_slice(d)
        exact[s] = t0 + flux / ks[d] * (xs[s] - cuts[d])
        t0 += flux * res[d]
    assert np.abs(x - exact).max() <= 1e-10


def test_itf_steady_driver_and_transient_steps(ctx):
    """SteadyBase::solve and three TransientBase::solve steps over the joined system, against the
    oracle's stepping (SuperLU per Picard iteration)."""
    print("[interface drivers]", flush=True)
    k, prd, q = _field_fns()
    p = itf_heat(8, splits=[3], k=k, prd=prd, q=q)
    p.to_device(ctx)
    st = ctx.steady_solve(p.ops, 20, 1e-12, 1.0, lin=dict(tol=1e-13))
    T = ctx.scldom_get(p.unknown)
    # the oracle's Picard iteration (the production's lagged skew term reads the unknown)
    x_lu = np.zeros(p.N)
    for _ in range(20):
        rp, col, val, b, _ = p.oracle_system(unknown_value=x_lu)
        x_new = spla.spsolve(sp.csr_matrix((val, col, rp), shape=(p.N, p.N)).tocsc(), b)
        done = np.abs(x_new - x_lu).max() <= 1e-14 * np.abs(x_new).max()
        x_lu = x_new
        if done:
            break
    assert np.linalg.norm(T - x_lu) / np.linalg.norm(x_lu) <= 1e-10, st.picard_iterations

    # transient: time operators on every row class, three implicit-Euler steps from T = 0
    w = p.scldom(vfvm.FIELD_CONSTANT, [1.0, 3.0])
    ops = []
    for o in p.ops:
        if o[0] == vfvm.OP_SCL_DOM_DIFFUSION:
            ops.append((vfvm.OP_SCL_DOM_TIME, o[1], o[2], w, -1, -1))
        elif o[0] == vfvm.OP_SCL_BND_DIFFUSION:
            ops.append((vfvm.OP_SCL_BND_TIME, o[1], o[2], w, -1, -1))
        elif o[0] == vfvm.OP_SCL_ITF_CONTINUITY:
            ops.append((vfvm.OP_SCL_ITF_TIME, o[1], o[2], w, -1, -1))
        ops.append(o)
    p.ops = ops
    p.to_device(ctx)
    dt = 2e-2
    ctx.transient_solve(p.ops, dt, 3, 10, 1e-12, 1.0, lin=dict(tol=1e-13))
    Tg = ctx.scldom_get(p.unknown)
    u = np.zeros(p.N)
    for (j, vals) in p.sbnd:
        d, lb = p.lbnd[j]
        if any(o[0] == vfvm.OP_SCL_BND_DIRICHLET and o[1] == j for o in p.ops):
            u[int(p.dom_ptr[d]) + p.parts[d].num_pts + p.parts[d].bnd_nodes(lb)] = vals
    for _ in range(3):
        prev = u.copy()
        for _it in range(20):
            rp, col, val, b, _ = p.oracle_system(dt=dt, value_prev={p.unknown: prev}, unknown_value=u)
            u_new = spla.spsolve(sp.csr_matrix((val, col, rp), shape=(p.N, p.N)).tocsc(), b)
            done = np.abs(u_new - u).max() <= 1e-14 * np.abs(u_new).max()
            u = u_new
            if done:
                break
    assert np.linalg.norm(Tg - u) / np.linalg.norm(u) <= 1e-9


def test_itf_rejections(ctx):
    """Flow operators, the adjoint and grad_coeff are outside the built path on joined meshes."""
    p = itf_heat(7)
    p.to_device(ctx)
    with pytest.raises(vfvm.VFVChemError) as e:
        p.mesh.build_grad_coeff()
    assert e.value.code == vfvm._lib.ERR_OPERATOR_UNSUPPORTED
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.adjoint_forward(p.ops)
    assert e.value.code == vfvm._lib.ERR_OPERATOR_UNSUPPORTED
    # an interface operator naming an interface that does not exist
    bad = list(p.ops) + [(vfvm.OP_SCL_ITF_SOURCE, 5, p.unknown, p.fields["src"], -1, -1)]
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.pattern_build(bad)
    assert e.value.code == vfvm._lib.ERR_INVALID_ARGUMENT
