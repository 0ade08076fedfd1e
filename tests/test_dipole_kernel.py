"""The built-in dipolar tensor kernel (HMX_KERNEL_DIPOLE, dims = 3), CPU side.

Reference: dipolar_tensor of examples/hmatrix_vector.jl:12-51 (a user matrix
there; built in here).  The oracle's restatement is pinned by closed forms and
exact identities; the host build with the built-in kernel must equal the
oracle's build bit for bit, and the Python callback of the same formula must
give the same layout as the built-in kernel.
"""
import math

import numpy as np
import pytest

import hmatrixgpu as H
import oracle_lib as O


def _dip(a, b, r):
    return O.lib.orc_dipole(a, b, float(r[0]), float(r[1]), float(r[2]))


def test_oracle_dipole_closed_forms():
    # r = (1, 0, 0): N = diag(-2, 1, 1) / (4 pi)
    f = 1 / (4 * math.pi)
    assert _dip(0, 0, (1, 0, 0)) == -2 / (4 * math.pi) and abs(_dip(0, 0, (1, 0, 0)) + 2 * f) == 0
    assert _dip(1, 1, (1, 0, 0)) == 1 / (4 * math.pi) and _dip(2, 2, (1, 0, 0)) == 1 / (4 * math.pi)
    assert _dip(0, 1, (1, 0, 0)) == 0.0 and _dip(0, 2, (1, 0, 0)) == 0.0
    # r = (1, 1, 0)/sqrt(2) scaled: xy = -3 x y / |r|^5 / 4 pi
    r = (1.0, 1.0, 0.0)
    assert math.isclose(_dip(0, 1, r), -3 / (2 ** 2.5) / (4 * math.pi), rel_tol=1e-15)
    assert all(_dip(a, b, (0, 0, 0)) == 0.0 for a in range(3) for b in range(3))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_dipole_identities(seed):
    rng = np.random.default_rng(seed)
    for r in rng.standard_normal((50, 3)):
        N = np.array([[_dip(a, b, r) for b in range(3)] for a in range(3)])
        assert np.array_equal(N, N.T)                              # symmetric, bitwise
        assert abs(np.trace(N)) <= 1e-14 * np.abs(N).max()          # trace-free
        d = np.linalg.norm(r)
        ref = -(3 * np.outer(r, r) - d * d * np.eye(3)) / (4 * math.pi * d ** 5)
        assert np.allclose(N, ref, rtol=1e-13, atol=1e-15 * np.abs(ref).max())
        # homogeneous of degree -3: exact under scaling by 2
        N2 = np.array([[_dip(a, b, 2 * r) for b in range(3)] for a in range(3)])
        assert np.array_equal(N2 * 8, N)


def _circle_points(n, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, 3)).T.copy()


def test_host_build_bit_exact_with_oracle():
    n = 300
    pts = _circle_points(n, 4)
    t_o = O.Tree(pts.T, 16, dims=3)
    t_p = H.ClusterTree(pts, max_points_per_leaf=16, dims=3)
    K = O.make_kernel(O.K_DIPOLE, xt=pts.T, xs=pts.T, m=3 * n, n=3 * n)
    Ho = O.HMatrix(K, t_o, t_o, 1.0, 1e-6)
    Hp = H.HMatrix(H.DipoleKernel(pts), t_p, t_p, eta=1.0, eps=1e-6, storage=())
    lay = Hp.layout(True)
    assert np.array_equal(lay["target_index_map"], Ho.target_index_map)
    assert np.array_equal(lay["dense_block_indices"], Ho.dense_block_indices.T)
    assert np.array_equal(lay["U_block_indices"], Ho.U_block_indices.T)
    assert np.array_equal(lay["dense_blocks"], Ho.dense_blocks)
    assert np.array_equal(lay["U_matrices"], Ho.U_matrices)
    assert np.array_equal(lay["V_matrices"], Ho.V_matrices)


def test_callback_of_the_same_formula_gives_the_same_layout():
    n = 200
    pts = _circle_points(n, 5)
    T = H.ClusterTree(pts, max_points_per_leaf=16, dims=3)
    Kb = H.DipoleKernel(pts)

    class Callback:                      # a user matrix with the same entries
        shape = Kb.shape

        def __getitem__(self, idx):
            return Kb[idx]

    a = H.HMatrix(Kb, T, T, eta=1.0, eps=1e-6, storage=()).layout(True)
    b = H.HMatrix(Callback(), T, T, eta=1.0, eps=1e-6, storage=()).layout(True)
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_dipole_needs_dims3():
    pts = _circle_points(50, 6)
    T = H.ClusterTree(pts, max_points_per_leaf=16)
    with pytest.raises(H.HMXError, match="dims == 3"):
        H.HMatrix(H.DipoleKernel(pts), T, T, eta=1.0, eps=1e-6, storage=())
