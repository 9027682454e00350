"""Parity of the HIP path (through the C ABI) with the oracle, on the GPU.

Bars: CSR pattern identical; assembled values and b equal to the oracle's to
1e-14 relative (the kernels follow the oracle's summation order, so they are
expected to be bitwise equal); FP64 GMRES solutions within 1e-10 relative L2 of
the FP64 sparse-LU solution (scipy SuperLU on the oracle's matrix), refined
mixed precision within 1e-8.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import vfvm
import vfvm_oracle as O
from problems import heat_c1, mass_c2, random_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


def _assembled(ctx, p, dt=0.0):
    p.to_device(ctx)
    ctx.pattern_build(p.ops)
    ctx.assemble(p.ops, dt)
    rp, col = ctx.pattern_get()
    vals, b = ctx.matrix_get()
    return rp, col, vals, b


def _lu(rp, col, val, b):
    A = sp.csr_matrix((val, col, rp), shape=(len(rp) - 1, len(rp) - 1)).tocsc()
    return spla.spsolve(A, b)


@pytest.mark.parametrize("n", [6, 12, 20])
def test_heat_pattern_values(ctx, n):
    print(f"[heat assemble n={n}]", flush=True)
    p = heat_c1(vfvm.Mesh.lattice(n))
    rp, col, vals, b = _assembled(ctx, p)
    orp, ocol, oval, ob, _ = p.oracle_system()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    scale = np.abs(oval).max()
    assert np.abs(vals - oval).max() <= 1e-14 * scale
    assert np.abs(b - ob).max() <= 1e-14 * max(np.abs(ob).max(), 1.0)
    print(" bitwise-equal values:", np.array_equal(vals, oval), "b:", np.array_equal(b, ob))


@pytest.mark.parametrize("n,c_out", [(8, None), (12, 0.0), (16, 0.5)])
def test_mass_pattern_values(ctx, n, c_out):
    print(f"[mass assemble n={n}]", flush=True)
    p = mass_c2(vfvm.Mesh.lattice(n), c_out=c_out, prod=0.3)
    rp, col, vals, b = _assembled(ctx, p)
    orp, ocol, oval, ob, _ = p.oracle_system()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.abs(vals - oval).max() <= 1e-14 * np.abs(oval).max()
    assert np.abs(b - ob).max() <= 1e-14 * max(np.abs(ob).max(), 1.0)


@pytest.mark.parametrize("precision,precond,bar", [
    (vfvm.FP64, vfvm.PRECOND_JACOBI, 1e-10), (vfvm.FP64, vfvm.PRECOND_ILU0, 1e-10), (vfvm.FP64, vfvm.PRECOND_NONE, 1e-10),
    (vfvm.FP32_REFINED, vfvm.PRECOND_JACOBI, 1e-8), (vfvm.FP32_REFINED, vfvm.PRECOND_ILU0, 1e-8)])
def test_heat_solve_vs_lu(ctx, precision, precond, bar):
    print(f"[heat solve prec={precision} pc={precond}]", flush=True)
    p = heat_c1(vfvm.Mesh.lattice(20))
    rp, col, vals, b = _assembled(ctx, p)
    x_lu = _lu(rp, col, vals, b)
    x, st = ctx.solve(None, tol=1e-13, restart=50, max_outer=400, precision=precision, precond=precond)
    err = np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu)
    print(f" iters={st.iterations} refinements={st.refinements} rel_res={st.rel_residual:.2e} err_vs_LU={err:.2e}")
    assert st.rel_residual <= 1e-13
    assert err <= bar


@pytest.mark.parametrize("precision,precond", [(vfvm.FP64, vfvm.PRECOND_JACOBI), (vfvm.FP32_REFINED, vfvm.PRECOND_JACOBI),
                                               (vfvm.FP64, vfvm.PRECOND_ILU0), (vfvm.FP32_REFINED, vfvm.PRECOND_ILU0)])
def test_mass_solve_vs_lu(ctx, precision, precond):
    print(f"[mass solve prec={precision} pc={precond}]", flush=True)
    p = mass_c2(vfvm.Mesh.lattice(16), c_out=0.0)
    rp, col, vals, b = _assembled(ctx, p)
    x_lu = _lu(rp, col, vals, b)
    # the convection-dominated system is worse conditioned (kappa ~ 5e3): the relative
    # residual must reach 1e-14 for a 1e-10 solution error
    x, st = ctx.solve(None, tol=1e-14, restart=50, max_outer=400, precision=precision, precond=precond)
    err = np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu)
    print(f" iters={st.iterations} rel_res={st.rel_residual:.2e} err_vs_LU={err:.2e}")
    assert err <= (1e-10 if precision == vfvm.FP64 else 1e-8)


def test_run_twice_bitwise(ctx):
    print("[determinism]", flush=True)
    p = mass_c2(vfvm.Mesh.lattice(14), c_out=0.0)
    outs = []
    for _ in range(2):
        _, _, vals, b = _assembled(ctx, p)
        x, _ = ctx.solve(None, tol=1e-12, precision=vfvm.FP64)
        xm, _ = ctx.solve(None, tol=1e-12, precision=vfvm.FP32_REFINED)
        outs.append((vals.copy(), b.copy(), x, xm))
    for a, c in zip(outs[0], outs[1]):
        assert np.array_equal(a, c)


def test_steady_driver_heat_closed_form(ctx):
    """SteadyBase::solve on config 1 (20^3): matches the LU solution of the same
    system and approximates T = x + q x (1 - x) / (2k) (heat_analytic.rs:9)."""
    print("[steady driver heat]", flush=True)
    mesh = vfvm.Mesh.lattice(20)
    p = heat_c1(mesh)
    p.to_device(ctx)
    st = ctx.steady_solve(p.ops, 10, 1e-8, 1.0, lin=dict(tol=1e-13))
    T = ctx.scldom_get(p.unknown)
    orp, ocol, oval, ob, _ = p.oracle_system()
    x_lu = _lu(orp, ocol, oval, ob)
    assert np.linalg.norm(T - x_lu) / np.linalg.norm(x_lu) <= 1e-10
    x = mesh.get("pos")[:, 0]
    exact = x + x * (1 - x) / 2.0
    assert np.abs(T - exact).max() < 2e-2
    assert st.picard_iterations == 2


def test_transient_driver_heat(ctx):
    """TransientBase::solve: 5 implicit Euler steps against the oracle's stepping."""
    print("[transient driver heat]", flush=True)
    mesh = vfvm.Mesh.lattice(8)
    p = heat_c1(mesh, q=0.0)
    w = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    T = p.unknown
    ops = [(vfvm.OP_SCL_DOM_TIME, 0, T, w, -1)] + p.ops[:5]
    rest = p.ops[5:]
    for i in range(0, len(rest), 4):
        ops += [(vfvm.OP_SCL_BND_TIME, rest[i][1], T, w, -1)] + rest[i:i + 4]
    p.ops = ops
    p.to_device(ctx)
    dt = 1e-2
    ctx.transient_solve(p.ops, dt, 5, 10, 1e-10, 1.0, lin=dict(tol=1e-13))
    Tg = ctx.scldom_get(T)
    # oracle stepping
    g = None
    u = np.zeros(p.N)
    nid0 = mesh.bnd_nodes(0) + mesh.num_pts
    nid1 = mesh.bnd_nodes(1) + mesh.num_pts
    u[nid0] = 0.0
    u[nid1] = 1.0
    for _ in range(5):
        prev = u.copy()
        for _it in range(2):
            rp, col, val, b, g = p.oracle_system(g=g, value_prev={T: prev}, dt=dt, unknown_value=u)
            u = _lu(rp, col, val, b)
    assert np.linalg.norm(Tg - u) / np.linalg.norm(u) <= 1e-9


def test_c2_full_size_uniform_solution(ctx):
    """Config 2 at 100^3: u.n = 0 on every wall and c = 1 at the inlet, so the exact
    discrete solution is c = 1 (uniform c is in the null space of every row);
    FP64 and refined-mixed GMRES reach 1e-10 and recover it."""
    print("[c2 100^3]", flush=True)
    p = mass_c2(vfvm.Mesh.lattice(100))
    p.to_device(ctx)
    ctx.pattern_build(p.ops)
    ctx.assemble(p.ops)
    for prec in (vfvm.FP64, vfvm.FP32_REFINED):
        x, st = ctx.solve(None, tol=1e-10, restart=50, max_outer=400, precision=prec)
        print(f" prec={prec} iters={st.iterations} rel_res={st.rel_residual:.2e} t={st.seconds:.3f}s")
        assert st.rel_residual <= 1e-10
        # the cellular flow with one Dirichlet face is poorly conditioned (closed
        # streamlines): the 1e-10 residual bounds the error only to ~1e-4
        assert np.abs(x - 1.0).max() <= 1e-3


@pytest.mark.parametrize("prob,n", [("heat", 8), ("mass", 8), ("mass", 13)])
def test_ilu0_application_vs_oracle(ctx, prob, n):
    """M^-1 r of the device ILU(0) against the oracle's ILU(0) (Saad Alg. 10.4, IKJ) of the
    same matrix in the device's row order (tiled: tile colours, tiles, levels); the oracle's
    factors are pinned by (LU)_ij = A_ij on the pattern."""
    print(f"[ilu0 {prob} n={n}]", flush=True)
    mesh = vfvm.Mesh.lattice(n)
    p = heat_c1(mesh) if prob == "heat" else mass_c2(mesh, c_out=0.0, prod=0.3)
    rp, col, vals, b = _assembled(ctx, p)
    perm, cptr = ctx.ilu_order()
    tptr, level = ctx.ilu_tiles()
    n = len(rp) - 1
    assert np.array_equal(np.sort(perm), np.arange(n)) and cptr[0] == 0 and cptr[-1] == n
    A = sp.csr_matrix((vals, col, rp), shape=(n, n))
    # the launches' independence: rows of one colour are not coupled (in A or A^T), except, in
    # the tiled order, rows of one tile at different levels
    colour = np.empty(n, np.int64)
    for q in range(len(cptr) - 1):
        colour[perm[cptr[q]:cptr[q + 1]]] = q
    tile = np.zeros(n, np.int64)
    lev = np.zeros(n, np.int64)
    if len(tptr):
        assert tptr[0] == 0 and tptr[-1] == n and np.all(np.diff(tptr) > 0)
        tile[perm] = np.repeat(np.arange(len(tptr) - 1), np.diff(tptr))
        lev[perm] = level
        # tiles do not straddle tile colours
        assert np.all(np.isin(cptr, tptr))
    else:
        tile[:] = np.arange(n)
    coo = A.tocoo()
    off = coo.row != coo.col
    r, c = coo.row[off], coo.col[off]
    same_colour = colour[r] == colour[c]
    assert np.all(tile[r][same_colour] == tile[c][same_colour])
    assert np.all(lev[r][same_colour] != lev[c][same_colour])
    L, U = O.ilu0(A, perm)
    P = sp.csr_matrix((np.ones(n), (np.arange(n), perm)), shape=(n, n))
    Ap = (P @ A @ P.T).tocsr()
    LU = (L @ U).tocsr()
    mask = Ap.copy()
    mask.data[:] = 1.0
    assert np.abs((LU.multiply(mask) - Ap).data).max() <= 1e-12 * np.abs(Ap.data).max()
    r = np.random.default_rng(4).normal(size=n)
    z_ref = P.T @ spla.spsolve_triangular(U, spla.spsolve_triangular(L, P @ r, lower=True, unit_diagonal=True), lower=False)
    z = ctx.precond_apply(r, precond=vfvm.PRECOND_ILU0, precision=vfvm.FP64)
    assert np.abs(z - z_ref).max() <= 1e-12 * np.abs(z_ref).max()
    z32 = ctx.precond_apply(r, precond=vfvm.PRECOND_ILU0, precision=vfvm.FP32_REFINED)
    assert np.abs(z32 - z_ref).max() <= 1e-4 * np.abs(z_ref).max()
    zj = ctx.precond_apply(r, precond=vfvm.PRECOND_JACOBI)
    assert np.allclose(zj, r / A.diagonal(), rtol=1e-15, atol=0)


@pytest.mark.parametrize("prob,n", [("mass", 30), ("heat", 40)])
def test_ilu0_one_launch_sweeps_at_size(ctx, prob, n):
    """The tile sweeps of one application run in one launch, each tile waiting for the completion
    flags of the adjacent tiles it reads: on meshes of many tiles in four tile colours, M^-1 r
    against the oracle's ILU(0) (the sparse-row IKJ in the device's row order) and triangular
    solves, and bitwise equal over repeated applications."""
    print(f"[ilu0 one-launch sweeps {prob} n={n}]", flush=True)
    mesh = vfvm.Mesh.lattice(n)
    p = heat_c1(mesh) if prob == "heat" else mass_c2(mesh, c_out=0.0, prod=0.3)
    rp, col, vals, b = _assembled(ctx, p)
    perm, cptr = ctx.ilu_order()
    tptr, level = ctx.ilu_tiles()
    assert len(tptr) - 1 >= 50 and len(cptr) - 1 >= 3, (len(tptr), len(cptr))
    n = len(rp) - 1
    A = sp.csr_matrix((vals, col, rp), shape=(n, n))
    L, U = O.ilu0_rows(A, perm)
    P = sp.csr_matrix((np.ones(n), (np.arange(n), perm)), shape=(n, n))
    r = np.random.default_rng(6).normal(size=n)
    y = spla.spsolve_triangular(L.tocsr(), P @ r, lower=True, unit_diagonal=True)
    z_ref = P.T @ spla.spsolve_triangular(U.tocsr(), y, lower=False)
    zs = [ctx.precond_apply(r, precond=vfvm.PRECOND_ILU0, precision=vfvm.FP64) for _ in range(3)]
    assert np.abs(zs[0] - z_ref).max() <= 1e-12 * np.abs(z_ref).max()
    for z in zs[1:]:
        assert np.array_equal(z, zs[0])
    z32 = [ctx.precond_apply(r, precond=vfvm.PRECOND_ILU0, precision=vfvm.FP32_REFINED) for _ in range(3)]
    assert np.abs(z32[0] - z_ref).max() <= 1e-4 * np.abs(z_ref).max()
    for z in z32[1:]:
        assert np.array_equal(z, z32[0])


def test_ilu0_one_launch_sweeps_random_points(ctx):
    """The tiled order and the one-launch sweeps on a mesh of random points (irregular pencils,
    tile sizes, tile graph and levels): M^-1 r against the oracle's ILU(0) in the device's order,
    bitwise repeatable, and the ILU(0)-preconditioned solve against SuperLU."""
    print("[ilu0 one-launch sweeps, random points]", flush=True)
    mesh = random_cube(20000, 600, seed=3)
    p = heat_c1(mesh)
    rp, col, vals, b = _assembled(ctx, p)
    perm, cptr = ctx.ilu_order()
    tptr, level = ctx.ilu_tiles()
    assert len(tptr) - 1 >= 50 and len(cptr) - 1 >= 3, (len(tptr), len(cptr))
    n = len(rp) - 1
    A = sp.csr_matrix((vals, col, rp), shape=(n, n))
    L, U = O.ilu0_rows(A, perm)
    P = sp.csr_matrix((np.ones(n), (np.arange(n), perm)), shape=(n, n))
    r = np.random.default_rng(8).normal(size=n)
    y = spla.spsolve_triangular(L.tocsr(), P @ r, lower=True, unit_diagonal=True)
    z_ref = P.T @ spla.spsolve_triangular(U.tocsr(), y, lower=False)
    zs = [ctx.precond_apply(r, precond=vfvm.PRECOND_ILU0, precision=vfvm.FP64) for _ in range(3)]
    assert np.abs(zs[0] - z_ref).max() <= 1e-12 * np.abs(z_ref).max()
    for z in zs[1:]:
        assert np.array_equal(z, zs[0])
    x, st = ctx.solve(None, tol=1e-12, restart=50, max_outer=400, precond=vfvm.PRECOND_ILU0, precision=vfvm.FP64)
    x_lu = _lu(rp, col, vals, b)
    assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) <= 1e-10


def test_ilu0_sweep_wait_that_gives_up_stops_the_solve(ctx):
    """A one-launch sweep whose wait for an adjacent tile gives up writes NaN (forced from the
    host: every such wait gives up at once), so M^-1 r is not finite and the solve reports
    LinearGmres non-finite instead of hanging; with the default limit restored the same system
    solves again, bit for bit as before."""
    print("[ilu0 wait limit]", flush=True)
    p = mass_c2(vfvm.Mesh.lattice(24), c_out=0.0, prod=0.3)
    rp, col, vals, b = _assembled(ctx, p)
    n = len(rp) - 1
    kw = dict(tol=1e-10, restart=50, max_outer=400, precond=vfvm.PRECOND_ILU0, precision=vfvm.FP64)
    x0, _ = ctx.solve(None, **kw)
    tptr, _ = ctx.ilu_tiles()
    assert len(tptr) - 1 > 8
    before = ctx.ilu_wait_limit(-1)
    z = ctx.precond_apply(np.ones(n), precond=vfvm.PRECOND_ILU0, precision=vfvm.FP64)
    assert np.isnan(z).any()
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.solve(None, **kw)
    assert e.value.code == vfvm._lib.ERR_LINEAR_NON_FINITE
    lapsed = ctx.ilu_wait_limit(1 << 17)
    assert lapsed > before
    x1, _ = ctx.solve(None, **kw)
    assert np.array_equal(x1, x0)


@pytest.mark.parametrize("prec,pc", [(vfvm.FP64, vfvm.PRECOND_ILU0), (vfvm.FP32_REFINED, vfvm.PRECOND_ILU0),
                                     (vfvm.FP64, vfvm.PRECOND_JACOBI)])
def test_cycle_graphs_match_eager_launches(ctx, prec, pc):
    """Solves replayed from the captured cycle graphs (kernel timing off) equal, bit for bit,
    the same solves launched kernel by kernel (timing on), over repeated replays with eager
    work in between."""
    print(f"[cycle graphs prec={prec} pc={pc}]", flush=True)
    p = mass_c2(vfvm.Mesh.lattice(12), c_out=0.0, prod=0.3)
    _assembled(ctx, p)
    kw = dict(tol=1e-11, restart=30, max_outer=400, precond=pc, precision=prec)
    xs = []
    for mode in (0, 0, -1, 0, -1, 0):
        ctx.timing_enable(mode)
        x, st = ctx.solve(None, **kw)
        assert st.rel_residual <= 1e-11
        xs.append(x)
    ctx.timing_enable(0)
    for x in xs[1:]:
        assert np.array_equal(x, xs[0])


def test_reference_api_mirror_steady_heat():
    """The host mirror of the reference API (Variables / SteadyHeat / LinearGmres) solves
    config 1 through the C ABI and matches the oracle's LU solution."""
    print("[api mirror]", flush=True)
    vars = vfvm.Variables.new_lattice(12)
    dom = vars.add_dom()
    b0, b1 = vars.add_bnd(dom, "x0"), vars.add_bnd(dom, "x1")
    T = vars.add_scldom_unkn(dom, 0.0)
    k = vars.add_scldom_cnst(dom, 1.0)
    zero = vars.add_scldom_cnst(dom, 0.0)
    q = vars.add_scldom_cnst(dom, 1.0)
    t0, t1 = vars.add_sclbnd_cnst(b0, 0.0), vars.add_sclbnd_cnst(b1, 1.0)
    phys = vfvm.SteadyHeat.new()
    phys.add_heat_dom(dom, T, k, zero, q)
    phys.add_temp_bnd(b0, t0)
    phys.add_temp_bnd(b1, t1)
    for name in ("y0", "y1", "z0", "z1"):
        b = vars.add_bnd(dom, name)
        phys.add_hflx_bnd(b, vars.add_sclbnd_cnst(b, 0.0))
    lin = vfvm.LinearGmres.new_with_options(1, 1e-13, 50, 200, precond=vfvm.PRECOND_ILU0)
    phys.solve(vars, lin, 10, 1e-8, 1.0)
    Tg = vars.value(T)
    p = heat_c1(vars.mesh)
    orp, ocol, oval, ob, _ = p.oracle_system()
    x_lu = _lu(orp, ocol, oval, ob)
    assert np.linalg.norm(Tg - x_lu) / np.linalg.norm(x_lu) <= 1e-10
    # error behaviour of SteadyBase::solve (steady_base.rs:21-23)
    with pytest.raises(vfvm.VFVChemError) as e:
        phys.solve(vars, lin, 0, 1e-8, 1.0)
    assert e.value.code == 40
    with pytest.raises(vfvm.VFVChemError) as e:
        phys.solve(vars, lin, 10, 1e-8, 1.5)
    assert e.value.code == 42


def test_linear_did_not_converge_is_reported(ctx):
    """GMRES failure maps to LinearGmresDidNotConverge (linear_gmres.rs:81)."""
    p = heat_c1(vfvm.Mesh.lattice(10))
    _assembled(ctx, p)
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.solve(None, tol=1e-14, restart=2, max_outer=1)
    assert e.value.code == 33
