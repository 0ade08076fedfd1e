"""sparsify_hmatrix (hmatrix.jl:234-274) through the C ABI, CPU only.

The expected matrix is assembled from the oracle's own block tree (merged, in
traverse order) and the oracle's kernel entries; the library's CSR must hold
exactly those entries (indices and values bitwise).  As in test_hmatrix.jl:106,
S*x also equals the dense-block part of the H-matrix product.
"""
import numpy as np
import scipy.sparse as sp

import hmatrixgpu as H
import oracle_lib as O
from helpers import uniform_points


def _expected(pts, leaf, eta):
    t = O.Tree(pts.T, leaf)
    b = O.BlockTree(t, t, eta, merge=True)
    direct, _ = b.traverse()
    K = O.make_kernel(O.K_LAPLACE, xt=pts.T, xs=pts.T)
    rows, cols, vals = [], [], []
    for tn, sn in direct:
        r = t.index_map[t.start[tn] - 1:t.end[tn] - 1] - 1
        c = t.index_map[t.start[sn] - 1:t.end[sn] - 1] - 1
        blk = O.kernel_matrix(K, r, c)
        rr, cc = np.meshgrid(r, c, indexing="ij")
        rows.append(rr.ravel())
        cols.append(cc.ravel())
        vals.append(blk.ravel())
    n = pts.shape[1]
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n)).tocsr(), direct


def test_sparsify_matches_oracle_blocks():
    pts = uniform_points(1500, 31)
    S_ref, direct = _expected(pts, 32, 1.5)
    X = H.ClusterTree(pts, max_points_per_leaf=32)
    S = H.sparsify_hmatrix(H.LaplaceKernel(pts), X, X, eta=1.5)
    S_ref.sort_indices()
    assert S.shape == S_ref.shape and S.nnz == S_ref.nnz
    assert np.array_equal(S.indptr, S_ref.indptr)
    assert np.array_equal(S.indices, S_ref.indices)
    assert np.array_equal(S.data, S_ref.data)  # bitwise: same entry arithmetic
    # the dense part of the H-matrix: every row appears, the diagonal entry is 0
    assert np.all(np.diff(S.indptr) > 0)
    assert np.all(S.diagonal() == 0.0)


def test_sparsify_equals_dense_block_product():
    pts = uniform_points(1200, 32)
    X = H.ClusterTree(pts, max_points_per_leaf=32)
    hm = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=1.5, eps=1e-6, storage=(), flatten=False)
    x = np.random.default_rng(3).standard_normal(1200)
    # dense_multiply of test_hmatrix.jl: only the direct (non-fallback) dense blocks
    d_ranges, _ = hm._ranges()
    lay = hm.layout(True)
    tm, smp = lay["target_index_map"] - 1, lay["source_index_map"] - 1
    di = lay["dense_block_indices"]
    y = np.zeros(1200)
    nd_direct = len(O.BlockTree(O.Tree(pts.T, 32), O.Tree(pts.T, 32), 1.5, merge=True).traverse()[0])
    D = lay["dense_blocks"]
    for bidx in range(nd_direct):
        r0, r1, c0, c1, off = di[:, bidx]
        blk = D[off:off + (r1 - r0 + 1) * (c1 - c0 + 1)].reshape(r1 - r0 + 1, c1 - c0 + 1)
        y[tm[r0 - 1:r1]] += blk @ x[smp[c0 - 1:c1]]
    S = H.sparsify_hmatrix(H.LaplaceKernel(pts), X, X, eta=1.5)
    assert np.allclose(S @ x, y, atol=1e-12, rtol=0)


def test_sparsify_callback_and_vector_kernel():
    # callback path with dims = 3 on the source side (test_hmatrix_vector.jl kernel)
    N = 300
    th = np.arange(1, N + 1) * 2 * np.pi / N
    pts = np.stack([np.sin(th), np.cos(th), np.zeros(N)])

    class Vec:
        shape = (N, 3 * N)

        def __getitem__(self, idx):
            r, c = idx
            J, j = np.asarray(c) // 3, np.asarray(c) % 3
            R = pts[:, np.asarray(r)][:, :, None] - pts[:, J][:, None, :]
            d = np.sqrt((R ** 2).sum(0))
            comp = np.take_along_axis(R, j[None, None, :].repeat(len(r), 1), 0)[0]
            return np.where(d < 1e-12, 1.0, comp / np.where(d == 0, 1, d) ** 2)

    T = H.ClusterTree(pts, max_points_per_leaf=16)
    Sx = H.ClusterTree(pts, max_points_per_leaf=16, dims=3)
    S = H.sparsify_hmatrix(Vec(), T, Sx, eta=1.5)
    assert S.shape == (N, 3 * N)
    A = S.toarray()
    full = Vec()[np.arange(N), np.arange(3 * N)]
    mask = A != 0
    assert np.array_equal(A[mask], full[mask])
