"""LinearBase for caller-assembled systems (linear_base.rs:5-10, linear_gmres.rs:53-84) through
the C ABI: the oracle's triplet lists in the reference's push order (duplicates included) give
the oracle's CSR pattern and values bit for bit (duplicates summed in list order, as faer's
argsort gather), and the device GMRES solves them to the bars of the operator path against
SuperLU (1e-10 FP64, 1e-8 refined mixed precision)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import vfvm
import vfvm_oracle as O
from problems import heat_c1, mass_c2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


def _triplets(p):
    g = O.dual(p.pos, p.mesh.get("tets"), p.mesh.num_pts)
    scl, sb, flux, bnd_nid, ops = p.oracle_inputs(g)
    return O.assemble(g, ops, scl, sb, flux, bnd_nid, flow_den=p.flow_den)


@pytest.mark.parametrize("make,n", [(heat_c1, 8), (mass_c2, 10)])
def test_triplets_to_csr_bitwise(ctx, make, n):
    print(f"[linear base triplets {make.__name__} n={n}]", flush=True)
    p = make(vfvm.Mesh.lattice(n))
    r, c, v, b = _triplets(p)
    nnz = ctx.linear_update_sparse(p.N, r, c)
    ctx.linear_build_matrix(v)
    rp, col = ctx.pattern_get()
    vals, _ = ctx.matrix_get()
    orp, ocol, oval = O.coo_to_csr(r, c, v, p.N)
    assert nnz == len(ocol)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(vals, oval)


@pytest.mark.parametrize("precision,precond,bar", [
    (vfvm.FP64, vfvm.PRECOND_JACOBI, 1e-10), (vfvm.FP64, vfvm.PRECOND_ILU0, 1e-10),
    (vfvm.FP32_REFINED, vfvm.PRECOND_ILU0, 1e-8), (vfvm.FP32_REFINED, vfvm.PRECOND_JACOBI, 1e-8)])
def test_solve_matrix_vs_lu(ctx, precision, precond, bar):
    print(f"[linear base solve prec={precision} pc={precond}]", flush=True)
    p = heat_c1(vfvm.Mesh.lattice(14), k=1.0, q=1.0)
    r, c, v, b = _triplets(p)
    # shuffle the list and split every value in two: the matrix is the same
    rng = np.random.default_rng(4)
    order = rng.permutation(r.size)
    r, c, v = r[order], c[order], v[order]
    half = 0.5 * v
    r, c, v = np.r_[r, r], np.r_[c, c], np.r_[half, v - half]
    ctx.linear_update_sparse(p.N, r, c)
    ctx.linear_build_matrix(v)
    A = sp.csr_matrix((v, (r, c)), shape=(p.N, p.N)).tocsc()
    xl = spla.spsolve(A, b)
    x0 = rng.standard_normal(p.N) * 0.1
    x, st = ctx.linear_solve_matrix(b, x0, tol=1e-13, restart=50, max_outer=1000, precision=precision, precond=precond)
    assert st.rel_residual <= 1e-13
    assert np.linalg.norm(x - xl) <= bar * np.linalg.norm(xl)
    # solve again with another right-hand side on the same matrix
    b2 = rng.standard_normal(p.N)
    x2, _ = ctx.linear_solve_matrix(b2, None, tol=1e-13, restart=50, max_outer=1000, precision=precision, precond=precond)
    assert np.linalg.norm(x2 - spla.spsolve(A, b2)) <= bar * np.linalg.norm(spla.spsolve(A, b2))


@pytest.mark.parametrize("per_row", [6, 12, 17, 19, 22, 27])
def test_solve_matrix_spmv_chunk_widths(ctx, per_row):
    """Random diagonally dominant matrices whose 128-row chunks need 8, 9, 10 or 12 column /
    value loads per thread of the block SpMV (6-22 entries per row; 27 per row takes the
    per-row kernel), 41 full chunks and a ragged last one (an odd count: the last workgroup
    takes one chunk): FP64 and refined mixed precision against SuperLU."""
    print(f"[linear base spmv chunk widths per_row={per_row}]", flush=True)
    n = 128 * 41 + 37
    rng = np.random.default_rng(per_row)
    k = per_row - 1
    r = np.repeat(np.arange(n), k)
    c = rng.integers(0, n, size=n * k)
    v = rng.uniform(-1.0, 1.0, size=n * k)
    d = np.bincount(r, weights=np.abs(v), minlength=n) + 1.0
    rows, cols, vals = np.r_[r, np.arange(n)], np.r_[c, np.arange(n)], np.r_[v, d]
    ctx.linear_update_sparse(n, rows, cols)
    ctx.linear_build_matrix(vals)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n)).tocsc()
    b = rng.standard_normal(n)
    xl = spla.spsolve(A, b)
    for precision, bar in ((vfvm.FP64, 1e-10), (vfvm.FP32_REFINED, 1e-8)):
        x, st = ctx.linear_solve_matrix(b, None, tol=1e-13, restart=50, max_outer=200, precision=precision,
                                        precond=vfvm.PRECOND_JACOBI)
        assert st.rel_residual <= 1e-13
        assert np.linalg.norm(x - xl) <= bar * np.linalg.norm(xl)


def test_linear_base_errors(ctx):
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.linear_update_sparse(3, [0, 1, 3], [0, 1, 2])  # row 3 of a 3 x 3 matrix
    assert e.value.code == 1
    ctx.linear_update_sparse(3, [0, 1, 2, 0], [0, 1, 2, 2])
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.linear_build_matrix([1.0, 2.0, 3.0])  # three values for four triplets
    assert e.value.code == 23
    ctx.linear_build_matrix([2.0, 3.0, 4.0, 1.0])
    x, _ = ctx.linear_solve_matrix(np.array([3.0, 3.0, 4.0]), None, tol=1e-14)
    assert np.allclose(x, [1.0, 1.0, 1.0], rtol=0, atol=1e-13)
    # the operator assembly refuses a system that came from triplets
    p = heat_c1(vfvm.Mesh.lattice(4))
    p.to_device(ctx)
    ctx.linear_update_sparse(2, [0, 1], [0, 1])
    ctx.linear_build_matrix([1.0, 1.0])
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.assemble(p.ops)
    assert e.value.code == 23


def test_mirror_linear_gmres():
    """LinearGmres.update_sparse / build_matrix / solve_matrix of the API mirror."""
    p = heat_c1(vfvm.Mesh.lattice(7))
    r, c, v, b = _triplets(p)
    lin = vfvm.LinearGmres.new_with_options(1, 1e-12, 50, 200, precond=vfvm.PRECOND_ILU0)
    lin.update_sparse((r, c, v), p.N)
    a = lin.build_matrix((r, c, v))
    x = lin.solve_matrix(a, b, np.zeros(p.N), p.N)
    xl = spla.spsolve(sp.csr_matrix((v, (r, c)), shape=(p.N, p.N)).tocsc(), b)
    assert np.linalg.norm(x - xl) <= 1e-10 * np.linalg.norm(xl)
