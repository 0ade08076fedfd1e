"""GPU parity of the batched product of uneven H-matrices (hmx_batch_* through the C ABI).

The batch concatenates the members' work lists and replays them as one captured
hipGraph; every work item computes exactly what it computes in the member's own
hmx_matvec, so at the same RHS width the batched result must be bit-identical to
the per-member product.  Each member is also checked against the oracle's
H-matvec (mult.jl:6-32 restated) on its exported layout:
  FP64 storage ... 1e-12     FP32 storage ... 1e-5   (relative 2-norm error)
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from helpers import box_instances, layout_ns, rel

pytestmark = pytest.mark.gpu

LEAF = 32


@pytest.fixture(scope="module")
def members(cuda):
    import hmatrixgpu as H
    pts_list = box_instances(6, 500, 6000, 61)
    print("[batch] building sizes", [p.shape[1] for p in pts_list], flush=True)
    hms = []
    for pts in pts_list:
        X = H.ClusterTree(pts, max_points_per_leaf=LEAF)
        hms.append(H.HMatrix(H.LaplaceKernel(pts), X, X, eta=2.0, eps=1e-6,
                             storage=("f64", "f32")))
    return hms, [layout_ns(h) for h in hms]


def _operands(torch, hms, nrhs, dtype, dev, seed):
    rng = np.random.default_rng(seed)
    Xs, Ys, Xn = [], [], []
    for h in hms:
        x = rng.standard_normal((h.n, nrhs))
        Xn.append(x)
        Xs.append(torch.tensor(x, dtype=dtype, device=dev).t().contiguous().t())
        Ys.append(torch.full((nrhs, h.m), float("nan"), dtype=dtype, device=dev).t())
    return Xs, Ys, Xn


def _singles(torch, H, hms, Xs, storage):
    out = []
    for h, x in zip(hms, Xs):
        y = torch.empty((x.shape[1], h.m), dtype=x.dtype, device=x.device).t()
        H.mul_(y, h, x, storage=storage)
        out.append(y)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("storage,dtype,tol", [("f32", "float32", 1e-5), ("f64", "float64", 1e-12)])
@pytest.mark.parametrize("nrhs", [1, 3, 4, 16])
def test_batch_matches_members_and_oracle(members, cuda, storage, dtype, tol, nrhs):
    import torch
    import hmatrixgpu as H
    hms, lays = members
    dt = getattr(torch, dtype)
    print("[batch] storage=%s nrhs=%d" % (storage, nrhs), flush=True)
    Xs, Ys, Xn = _operands(torch, hms, nrhs, dt, cuda, 100 + nrhs)
    B = H.HMatrixBatch(hms, storage=storage, nrhs=nrhs, io=dtype)
    B.matvec(Ys, Xs)
    torch.cuda.synchronize()
    ref = _singles(torch, H, hms, Xs, storage)
    for i, (y, r) in enumerate(zip(Ys, ref)):
        assert torch.equal(y, r), "member %d differs from its own product" % i
        Yo = np.stack([O.mul_layout(lays[i], Xn[i][:, c]) for c in range(nrhs)], axis=1)
        assert rel(y.cpu().numpy(), Yo) <= tol


def test_graph_replays_with_eager_work_between(members, cuda):
    import torch
    import hmatrixgpu as H
    hms, lays = members
    nrhs = 4
    Xs, Ys, _ = _operands(torch, hms, nrhs, torch.float32, cuda, 7)
    B = H.HMatrixBatch(hms, storage="f32", nrhs=nrhs)
    B.matvec(Ys, Xs)
    first = [y.clone() for y in Ys]
    for rep in range(4):
        # eager work on the same stream between replays: new X contents in place,
        # an unrelated reduction, a host synchronisation
        g = torch.Generator(device=cuda).manual_seed(rep)
        for x in Xs:
            x.copy_(torch.randn(x.shape, generator=g, device=cuda, dtype=torch.float32))
        _ = torch.randn(1 << 20, device=cuda).sum().item()
        B.matvec(Ys, Xs)
        ref = _singles(torch, H, hms, Xs, "f32")
        for i, (y, r) in enumerate(zip(Ys, ref)):
            assert torch.equal(y, r), "replay %d member %d" % (rep, i)
    # new operand buffers: the descriptor table is refreshed before the replay
    Xs2 = [x.clone() for x in Xs]
    Ys2 = [torch.zeros_like(y) for y in Ys]
    for x in Xs:
        x.fill_(0)
    B.matvec(Ys2, Xs2)
    ref = _singles(torch, H, hms, Xs2, "f32")
    for y, r in zip(Ys2, ref):
        assert torch.equal(y, r)
    assert any(not torch.equal(a, b) for a, b in zip(first, Ys2))


def test_autotune_picks_valid_config_and_caches(members, cuda, tmp_path):
    import torch
    import hmatrixgpu as H
    hms, _ = members
    nrhs = 4
    cache = tmp_path / "tune.txt"
    os.environ["HMX_TUNE_CACHE"] = str(cache)
    try:
        Xs, Ys, _ = _operands(torch, hms, nrhs, torch.float32, cuda, 9)
        B = H.HMatrixBatch(hms, storage="f32", nrhs=nrhs)
        assert B.config() == (4, "lpt")
        ms = B.autotune(Ys, Xs, iters=3)
        assert ms > 0
        w, order = B.config()
        assert w in (4, 16) and order in H.HMatrixBatch.ORDERS
        assert cache.read_text().split()[1:] == [str(w), str(H.HMatrixBatch.ORDERS.index(order))]
        B.matvec(Ys, Xs)
        torch.cuda.synchronize()
        # the tuned batch still agrees with the member products to FP32 accuracy
        ref = _singles(torch, H, hms, Xs, "f32")
        for y, r in zip(Ys, ref):
            assert rel(y.cpu().numpy(), r.cpu().numpy()) <= 1e-5
        # a new batch of the same shape starts from the cached choice
        B2 = H.HMatrixBatch(hms, storage="f32", nrhs=nrhs)
        assert B2.config() == (w, order)
    finally:
        del os.environ["HMX_TUNE_CACHE"]


def test_forced_widths_agree(members, cuda):
    import torch
    import hmatrixgpu as H
    hms, _ = members
    Xs, Ys, _ = _operands(torch, hms, 2, torch.float32, cuda, 13)
    B = H.HMatrixBatch(hms, storage="f32", nrhs=2)
    out = {}
    for w in (4, 16):
        for order in H.HMatrixBatch.ORDERS:
            B.set_config(w, order)
            B.matvec(Ys, Xs)
            torch.cuda.synchronize()
            out[(w, order)] = [y.cpu().numpy().copy() for y in Ys]
    base = out[(4, "lpt")]
    for key, ys in out.items():
        for a, b in zip(ys, base):
            if key[0] == 4:
                assert np.array_equal(a, b), key   # order changes only the schedule
            else:
                assert rel(a, b) <= 1e-5, key


def test_members_are_pinned_while_the_batch_lives(members, cuda):
    import hmatrixgpu as H
    hms, _ = members
    B = H.HMatrixBatch(hms[:2], storage="f32", nrhs=4)
    with pytest.raises(H.HMXError, match="live batch"):
        hms[0].release("f32")
    with pytest.raises(H.HMXError, match="live batch"):
        hms[1].materialize("f32")
    del B
    import gc
    gc.collect()
    hms[0].materialize("f32")   # allowed again once the batch is gone
