"""Config 2 at its full size: N = 65,536 uniform [0,1)^3, 1/r, leaf 64, eta 2,
tol 1e-6, one N(0,1) right-hand side in FP64 -- the drop-in mul!.

Reference behaviour: mul!(result, H, x) (mult.jl:159-189) equals the CPU product
over the same blocks (mult.jl:6-32).  The checker is the oracle's restatement
of that product evaluated on the layout the build exports:
  FP64 storage, 1 RHS ..... relative 2-norm error <= 1e-12
  FP32 storage, 16 RHS .... relative 2-norm error <= 1e-5
and the GPU ACA+ build of the same matrix has the identical layout (bitwise).
"""
import numpy as np
import pytest

import oracle_lib as O
from helpers import layout_ns, rel

pytestmark = pytest.mark.gpu

N = 65536


@pytest.fixture(scope="module")
def c2(cuda):
    import hmatrixgpu as H
    pts = np.random.default_rng(2).random((N, 3)).T.copy()
    X = H.ClusterTree(pts, max_points_per_leaf=64)
    print("[c2] host build N=%d" % N, flush=True)
    hm = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=2.0, eps=1e-6, storage=("f64", "f32"))
    return pts, X, hm, layout_ns(hm)


def test_fp64_single_rhs_config2(c2, cuda):
    import torch
    import hmatrixgpu as H
    _, _, hm, lay = c2
    x = np.random.default_rng(22).standard_normal(N)
    y_ref = O.mul_layout(lay, x)
    y = H.mul_(torch.empty(N, dtype=torch.float64, device=cuda), hm,
               torch.tensor(x, device=cuda), storage="f64")
    torch.cuda.synchronize()
    err = rel(y.cpu().numpy(), y_ref)
    print("[c2] fp64 rel err %.3e" % err, flush=True)
    assert err <= 1e-12


def test_fp32_16rhs_config2(c2, cuda):
    import torch
    import hmatrixgpu as H
    _, _, hm, lay = c2
    Xn = np.random.default_rng(23).standard_normal((N, 16))
    Yref = np.stack([O.mul_layout(lay, Xn[:, c]) for c in range(16)], axis=1)
    Xt = torch.tensor(Xn, dtype=torch.float32, device=cuda).t().contiguous().t()
    Yt = torch.empty((16, N), dtype=torch.float32, device=cuda).t()
    H.mul_(Yt, hm, Xt, storage="f32")
    torch.cuda.synchronize()
    err = rel(Yt.cpu().numpy(), Yref)
    print("[c2] fp32 16-rhs rel err %.3e" % err, flush=True)
    assert err <= 1e-5


def test_gpu_aca_layout_identical_config2(c2, cuda):
    import hmatrixgpu as H
    pts, X, hm, lay = c2
    print("[c2] GPU ACA+ build", flush=True)
    dev = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=2.0, eps=1e-6, storage=(), aca="gpu")
    ld = layout_ns(dev)
    for k in ("target_index_map", "source_index_map", "dense_block_indices", "U_block_indices",
              "V_block_indices", "U_matrices", "V_matrices"):
        assert np.array_equal(getattr(lay, k), getattr(ld, k)), k
    assert np.array_equal(hm.block_features()["aca_rank"], dev.block_features()["aca_rank"])
