"""The oracle's ILU(0) pinned by its defining properties: LU = A on the pattern, and the
exact LU factorisation where ILU(0) has no dropped fill (a tridiagonal matrix)."""
import numpy as np
import scipy.sparse as sp

import vfvm_oracle as O


def test_ilu0_reproduces_pattern_entries():
    rng = np.random.default_rng(0)
    n = 60
    A = sp.random(n, n, density=0.08, random_state=1, format="csr")
    A = A + A.T + sp.diags(np.full(n, 10.0))  # symmetric pattern, dominant diagonal
    A = A.tocsr()
    perm = rng.permutation(n)
    L, U = O.ilu0(A, perm)
    P = sp.csr_matrix((np.ones(n), (np.arange(n), perm)), shape=(n, n))
    Ap = (P @ A @ P.T).tocsr()
    S = Ap.copy()
    S.data[:] = 1.0
    diff = (L @ U).multiply(S) - Ap
    assert np.abs(diff.data).max() <= 1e-13 * np.abs(Ap.data).max()
    assert np.allclose(L.diagonal(), 1.0) and sp.triu(L, 1).nnz == 0 and sp.tril(U, -1).nnz == 0
    # fill entries exist in LU outside the pattern (the dropped terms): ILU(0) is not exact here
    assert np.abs((L @ U - Ap).toarray()).max() > 1e-8


def test_ilu0_exact_without_fill():
    n = 40
    main = np.linspace(3.0, 5.0, n)
    A = sp.diags([np.full(n - 1, -1.2), main, np.full(n - 1, -0.7)], [-1, 0, 1], format="csr")
    L, U = O.ilu0(A, np.arange(n))
    assert np.abs((L @ U - A).toarray()).max() <= 1e-14


def test_ilu0_rows_equals_dense_ilu0():
    """The sparse-row restatement gives the dense ilu0's factors bit for bit (same updates, same
    order), on a convection-diffusion matrix of the lattice in a random row order."""
    from problems import advdiff_1d
    import vfvm
    p = advdiff_1d(vfvm.Mesh.lattice(6, 3, 0.3), 8.0)
    rp, col, val, b, _ = p.oracle_system()
    n = len(rp) - 1
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    perm = np.random.default_rng(5).permutation(n)
    L, U = O.ilu0(A, perm)
    L2, U2 = O.ilu0_rows(A, perm)
    assert (L != L2).nnz == 0 and (U != U2).nnz == 0
