"""GPU products of H-matrices read back from disk (hmx_hmatrix_load).

A loaded operator holds the same index maps, blocks and factors as the saved
one, so its device plan and its products are the same: bitwise equal to the
original's, in FP64 and FP32 storage, for a built-in kernel (entries generated
on the device from the saved coordinates) and for a user matrix (saved dense
payload).  The FP64 product is also checked against the oracle's H-matvec on
the loaded layout (relative 2-norm error <= 1e-12).
"""
import numpy as np
import pytest

import oracle_lib as O
from helpers import layout_ns, rel, uniform_points

pytestmark = pytest.mark.gpu

N = 20000


def _products(cuda, hm, X):
    import torch
    import hmatrixgpu as H
    out = {}
    for st, dt in (("f64", torch.float64), ("f32", torch.float32)):
        Xt = torch.tensor(X, dtype=dt, device=cuda).t().contiguous().t()
        Yt = torch.empty((X.shape[1], hm.m), dtype=dt, device=cuda).t()
        H.mul_(Yt, hm, Xt, storage=st)
        torch.cuda.synchronize()
        out[st] = Yt.cpu().numpy()
    return out


def test_loaded_builtin_kernel_products_are_bitwise_equal(cuda, tmp_path):
    import hmatrixgpu as H
    pts = uniform_points(N, 51)
    T = H.ClusterTree(pts, max_points_per_leaf=64)
    print("[serialize] build N=%d" % N, flush=True)
    hm = H.HMatrix(H.LaplaceKernel(pts), T, T, eta=2.0, eps=1e-6, storage=("f64", "f32"))
    p = tmp_path / "lap.hmx"
    hm.save(p)
    g = H.HMatrix.load(p, storage=("f64", "f32"))
    X = np.random.default_rng(52).standard_normal((N, 4))
    print("[serialize] products", flush=True)
    a, b = _products(cuda, hm, X), _products(cuda, g, X)
    for st in a:
        assert np.array_equal(a[st], b[st]), st
    lay = layout_ns(g)
    assert rel(b["f64"][:, 0], O.mul_layout(lay, X[:, 0])) <= 1e-12


def test_loaded_user_matrix_products_are_bitwise_equal(cuda, tmp_path):
    import hmatrixgpu as H
    n = 6000
    pts = uniform_points(n, 53)
    T = H.ClusterTree(pts, max_points_per_leaf=32)
    Kb = H.LaplaceKernel(pts)

    class Callback:
        shape = Kb.shape

        def __getitem__(self, idx):
            return Kb[idx]

    print("[serialize] callback build n=%d" % n, flush=True)
    hm = H.HMatrix(Callback(), T, T, eta=1.5, eps=1e-6, storage=("f64", "f32"))
    p = tmp_path / "cb.hmx"
    hm.save(p)
    g = H.HMatrix.load(p, storage=("f64", "f32"))
    X = np.random.default_rng(54).standard_normal((n, 3))
    a, b = _products(cuda, hm, X), _products(cuda, g, X)
    for st in a:
        assert np.array_equal(a[st], b[st]), st
