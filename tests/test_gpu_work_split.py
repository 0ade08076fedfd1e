"""GPU parity of the product's work splitting, forced onto small operators.

Leaf items are cut along k into pieces whose partial rows k_leaf_reduce sums,
and far-field source clusters wider than a column chunk write partial rows that
k_reduce sums.  Large operators split by themselves; here the planner's knobs
(read when an operator is materialised) force both onto H-matrices of a few
thousand points: HMX_LEAF_PIECES / HMX_LEAF_MIN_PIECE cut nearly every leaf
item into many pieces, HMX_FAR_CHUNK = 64 cuts every far-field cluster wider
than 64 columns.  Checker: the oracle's H-matvec (mult.jl:6-32 restated) on the
exported layout, FP64 <= 1e-12 and FP32 <= 1e-5 (relative 2-norm error); the
batched product over split members equals each member's own product bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as O
from helpers import box_instances, layout_ns, rel, uniform_points

pytestmark = pytest.mark.gpu

FORCE = {"HMX_LEAF_PIECES": "100000", "HMX_LEAF_MIN_PIECE": "4096", "HMX_FAR_CHUNK": "64"}


def _force(monkeypatch):
    for k, v in FORCE.items():
        monkeypatch.setenv(k, v)


def _apply(torch, H, hm, X, storage, dt, dev):
    Xt = torch.tensor(X, dtype=dt, device=dev).t().contiguous().t()
    Yt = torch.full((X.shape[1], hm.m), float("nan"), dtype=dt, device=dev).t()
    H.mul_(Yt, hm, Xt, storage=storage)
    torch.cuda.synchronize()
    return Yt.cpu().numpy()


def test_split_single_operator_matches_oracle(cuda, monkeypatch):
    import torch
    import hmatrixgpu as H
    n = 6000
    pts = uniform_points(n, 71)
    T = H.ClusterTree(pts, max_points_per_leaf=32)
    hm = H.HMatrix(H.LaplaceKernel(pts), T, T, eta=1.5, eps=1e-6, storage=())
    lay = layout_ns(hm)
    hm.materialize(("f64", "f32"))
    plain_bytes = hm.info()["device_index_bytes"]
    hm.release(("f64", "f32"))
    _force(monkeypatch)
    print("[split] forced materialise", flush=True)
    hm.materialize(("f64", "f32"))
    split_bytes = hm.info()["device_index_bytes"]
    assert split_bytes > plain_bytes  # leaf groups and the extra items are in the plan
    rng = np.random.default_rng(72)
    for nrhs in (1, 4, 16):
        X = rng.standard_normal((n, nrhs))
        Yo = np.stack([O.mul_layout(lay, X[:, c]) for c in range(nrhs)], axis=1)
        y64 = _apply(torch, H, hm, X, "f64", torch.float64, cuda)
        y32 = _apply(torch, H, hm, X, "f32", torch.float32, cuda)
        e64, e32 = rel(y64, Yo), rel(y32, Yo)
        print("[split] nrhs %d fp64 %.2e fp32 %.2e" % (nrhs, e64, e32), flush=True)
        assert e64 <= 1e-12 and e32 <= 1e-5


@pytest.mark.parametrize("storage,dtype,nrhs", [("f64", "float64", 16), ("f32", "float32", 4)])
def test_split_batch_equals_members(cuda, monkeypatch, storage, dtype, nrhs):
    import torch
    import hmatrixgpu as H
    _force(monkeypatch)
    hms = []
    for pts in box_instances(4, 1500, 5000, 73):
        X = H.ClusterTree(pts, max_points_per_leaf=32)
        hms.append(H.HMatrix(H.LaplaceKernel(pts), X, X, eta=2.0, eps=1e-6, storage=(storage,)))
    dt = getattr(torch, dtype)
    rng = np.random.default_rng(74)
    Xn = [rng.standard_normal((h.n, nrhs)) for h in hms]
    Xs = [torch.tensor(x, dtype=dt, device=cuda).t().contiguous().t() for x in Xn]
    Ys = [torch.full((nrhs, h.m), float("nan"), dtype=dt, device=cuda).t() for h in hms]
    print("[split] batch storage=%s nrhs=%d" % (storage, nrhs), flush=True)
    B = H.HMatrixBatch(hms, storage=storage, nrhs=nrhs, io=dtype)
    B.matvec(Ys, Xs)
    torch.cuda.synchronize()
    for i, h in enumerate(hms):
        y = torch.empty((nrhs, h.m), dtype=dt, device=cuda).t()
        H.mul_(y, h, Xs[i], storage=storage)
        torch.cuda.synchronize()
        assert torch.equal(Ys[i], y), "member %d differs from its own product" % i
        Yo = np.stack([O.mul_layout(layout_ns(h), Xn[i][:, c]) for c in range(nrhs)], axis=1)
        assert rel(y.cpu().numpy(), Yo) <= (1e-12 if storage == "f64" else 1e-5)


def test_dispatch_order_changes_nothing(cuda, monkeypatch):
    """Each leaf item's arithmetic is independent of when it runs, so every
    dispatch order of the leaf items (HMX_ITEM_ORDER) gives the same bits."""
    import torch
    import hmatrixgpu as H
    n = 6000
    pts = uniform_points(n, 75)
    T = H.ClusterTree(pts, max_points_per_leaf=32)
    hm = H.HMatrix(H.LaplaceKernel(pts), T, T, eta=1.5, eps=1e-6, storage=())
    X = np.random.default_rng(76).standard_normal((n, 16))
    _force(monkeypatch)
    ys = {}
    for order in ("xcd", "lpt", "tree"):
        monkeypatch.setenv("HMX_ITEM_ORDER", order)
        hm.materialize("f32")
        ys[order] = _apply(torch, H, hm, X, "f32", torch.float32, cuda)
        hm.release("f32")
    assert np.array_equal(ys["xcd"], ys["lpt"]) and np.array_equal(ys["xcd"], ys["tree"])
