"""GPU batched ACA+ (hmx_hmatrix_build_ex with aca_device = 1).

Reference behaviour: ACA_plus (aca.jl:36-106) called per admissible block by
build_matrices (hmatrix.jl:189-213).  The device ACA+ must reproduce the host
ACA+ -- itself bit-exact with the oracle's restatement -- bit for bit: same
pivots, same terms, same ranks, hence an identical flattened layout after the
recompression and storage test.  Rank capacities (initial_rank) only change the
schedule: a block that outgrows its capacity continues in a further pass with
the same factors.
"""
import numpy as np
import pytest

from helpers import box_instances, uniform_points

pytestmark = pytest.mark.gpu


def _layout(hm):
    lay = hm.layout(True)
    return {k: np.asarray(v) for k, v in lay.items()}


def _same(a, b):
    la, lb = _layout(a), _layout(b)
    assert la.keys() == lb.keys()
    for k in la:
        assert la[k].shape == lb[k].shape and np.array_equal(la[k], lb[k]), k
    fa, fb = a.block_features(), b.block_features()
    assert np.array_equal(fa["aca_rank"], fb["aca_rank"])
    assert np.array_equal(fa["rank"], fb["rank"])
    assert np.array_equal(fa["features"], fb["features"])


@pytest.fixture(scope="module")
def c1(cuda):
    import hmatrixgpu as H
    pts = uniform_points(4096, 1)
    X = H.ClusterTree(pts, max_points_per_leaf=32)
    print("[aca] host build N=4096", flush=True)
    host = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=2.0, eps=1e-6, storage=())
    return pts, X, host


@pytest.mark.parametrize("initial_rank", [4, 16, 64])
def test_gpu_aca_bit_identical_to_host(c1, cuda, initial_rank):
    import hmatrixgpu as H
    pts, X, host = c1
    print("[aca] gpu build initial_rank=%d" % initial_rank, flush=True)
    dev = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=2.0, eps=1e-6, storage=(), aca="gpu",
                    initial_rank=initial_rank)
    _same(host, dev)
    st = dev.build_stats
    f = dev.block_features()
    assert st["blocks"] == len(f["aca_rank"]) > 0
    assert st["used_elems"] == int(np.sum((f["aca_rank"].astype(np.int64)) *
                                          _block_mn(host)))
    assert st["slot_elems"] >= st["used_elems"]
    # a block continues past its first capacity iff it reached it without the
    # stop test being met: every block whose rank exceeds its capacity did, a
    # block whose rank equals a capacity below min(m, n) may have
    rank, cap, kmax = f["aca_rank"], f["capacity"], _kmax(host)
    assert np.all(cap == np.minimum(np.maximum(4, (initial_rank + 3) // 4 * 4), kmax))
    lo = int(np.sum(rank > cap))
    hi = int(np.sum((rank >= cap) & (cap < kmax)))
    print("[aca] stats", st, "bounds", lo, hi, flush=True)
    assert lo <= st["unconverged_first"] <= hi
    assert (st["passes"] > 1) == (st["unconverged_first"] > 0)
    if initial_rank == 4:
        assert st["passes"] > 1


def _block_mn(hm):
    f = hm.block_features()["features"]
    return np.rint(np.exp(f[:, 3]) + np.exp(f[:, 4])).astype(np.int64)


def _kmax(hm):
    f = hm.block_features()["features"]
    return np.minimum(np.rint(np.exp(f[:, 3])), np.rint(np.exp(f[:, 4]))).astype(np.int64)


@pytest.mark.parametrize("eps", [1e-4, 1e-8])
def test_gpu_aca_other_geometry_and_tolerance(cuda, eps):
    import hmatrixgpu as H
    pts = box_instances(1, 3000, 3000, 17)[0]
    X = H.ClusterTree(pts, max_points_per_leaf=32)
    print("[aca] box N=%d eps=%g" % (pts.shape[1], eps), flush=True)
    host = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=2.0, eps=eps, storage=())
    dev = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=2.0, eps=eps, storage=(), aca="gpu",
                    initial_rank=8)
    _same(host, dev)


def test_gpu_aca_rectangular_blocks(cuda):
    import hmatrixgpu as H
    rng = np.random.default_rng(8)
    xt = rng.random((3, 2500))
    xs = rng.random((3, 1700)) + np.array([[1.5], [0.0], [0.0]])
    T = H.ClusterTree(xt, max_points_per_leaf=32)
    S = H.ClusterTree(xs, max_points_per_leaf=32)
    host = H.HMatrix(H.LaplaceKernel(xt, xs), T, S, eta=2.0, eps=1e-6, storage=())
    dev = H.HMatrix(H.LaplaceKernel(xt, xs), T, S, eta=2.0, eps=1e-6, storage=(), aca="gpu")
    _same(host, dev)


def test_gpu_aca_then_product(c1, cuda):
    import torch
    import hmatrixgpu as H
    pts, X, host = c1
    dev = H.HMatrix(H.LaplaceKernel(pts), X, X, eta=2.0, eps=1e-6, storage=("f64",), aca="gpu")
    x = torch.tensor(np.random.default_rng(3).standard_normal(4096), device=cuda)
    host.materialize(("f64",))
    y1 = H.mul_(torch.empty_like(x), host, x, storage="f64")
    y2 = H.mul_(torch.empty_like(x), dev, x, storage="f64")
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)


def test_gpu_aca_rejects_callback_kernels(c1, cuda):
    import hmatrixgpu as H
    pts, X, _ = c1
    K = H.Log2DKernel(pts)
    with pytest.raises(H.HMXError, match="1/r"):
        H.HMatrix(K, X, X, eta=2.0, eps=1e-6, storage=(), aca="gpu")
