"""Pins for the oracle's ACA+ (aca.jl:9-106) and SVD recompression
(aca.jl:123-143).

Known answer from the reference (test_aca.jl:7-42): K[i,j] = N/(i-j)^2, N=100,
bottom-left 30x30 block, ACA at 1e-7 -> rank < 30, |UV - B|_2 < 1e-8 and
max|UVx - Bx| < 1e-8; recompression keeps the errors and does not grow the rank.
LAPACK (scipy gesdd / geqrf, the routines Julia's svd/qr call) pins the QR and
SVD building blocks; the Eckart-Young identity pins the truncation.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle_lib as O


def _mat_kernel(B):
    return O.make_kernel(O.K_MATRIX, mat=B)


def _aca(B, eps, max_iter=0):
    k = _mat_kernel(B)
    return O.aca_plus(k, np.arange(B.shape[0]), np.arange(B.shape[1]), eps, max_iter)


def _ref_block():
    N = 100
    i = np.arange(1, N + 1)[:, None]
    j = np.arange(1, N + 1)[None, :]
    with np.errstate(divide="ignore"):
        K = np.where(i != j, N / (i - j) ** 2.0, 0.0)
    return K[N - 30:, :30]


def test_reference_known_answer():
    B = _ref_block()
    eps = 1e-6
    U, V = _aca(B, eps / 10.0)
    x = np.random.default_rng(0).random(30)
    assert U.shape[1] < 30
    assert np.linalg.norm(U @ V - B, 2) < eps / 100
    assert np.max(np.abs(U @ (V @ x) - B @ x)) < eps / 100
    U2, V2 = O.svd_recompress(U, V, eps / 10)
    assert U2.shape[1] <= U.shape[1]
    assert np.linalg.norm(U2 @ V2 - B, 2) < eps / 100
    assert np.max(np.abs(U2 @ (V2 @ x) - B @ x)) < eps / 100


def test_first_step_takes_column_one():
    # k = 1: pivot (1,1) with a zero residual -> the else branch fetches column 1
    # unchanged, picks the row of largest |entry|, and divides that row by its
    # entry in column 1: the column pivot is NOT re-chosen in this branch
    # (aca.jl:69-70, 83-86, 92-93)
    B = np.random.default_rng(1).random((20, 15))
    U, V = _aca(B, 1e-30, max_iter=1)
    assert U.shape[1] == 1
    assert np.array_equal(U[:, 0], B[:, 0])
    r = int(np.argmax(np.abs(B[:, 0])))
    assert np.array_equal(V[0], B[r] / B[r, 0])


def test_exact_rank_one_and_two():
    rng = np.random.default_rng(2)
    u, v = rng.random(40), rng.random(30)
    B = np.outer(u, v)
    U, V = _aca(B, 1e-10)
    # the stop test is applied after the term is pushed: the second term of an
    # exact rank-1 matrix has (numerically) zero norm and stops the loop
    assert U.shape[1] in (1, 2)
    assert np.allclose(U @ V, B, rtol=0, atol=1e-14 * np.abs(B).max())
    W = rng.random((40, 2)) @ rng.random((2, 30))
    U, V = _aca(W, 1e-10)
    assert U.shape[1] in (2, 3)
    assert np.allclose(U @ V, W, rtol=0, atol=1e-13 * np.abs(W).max())


def test_stop_criterion_and_max_iter():
    B = _ref_block()
    U, V = _aca(B, 1e-7)
    k = U.shape[1]
    # the last pushed term is the first one below eps (aca.jl:95-99)
    last = np.sqrt(np.sum(U[:, -1] ** 2) * np.sum(V[-1] ** 2))
    assert last < 1e-7
    for q in range(k - 1):
        assert np.sqrt(np.sum(U[:, q] ** 2) * np.sum(V[q] ** 2)) >= 1e-7
    U, V = _aca(B, 0.0, max_iter=5)
    assert U.shape[1] == 5
    U, V = _aca(B[:4], 0.0)  # capped by min(m, n)
    assert U.shape[1] == 4


def test_pivot_ties_lowest_index():
    # Column 1 has equal magnitudes everywhere: the row pivot is the first row
    # (stable descending sortperm, aca.jl:10).
    B = np.array([[1.0, 3.0, -3.0], [-1.0, 2.0, 0.5], [1.0, 0.0, 1.0]])
    U, V = _aca(B, 1e-30, max_iter=1)
    assert np.array_equal(V[0], B[0] / B[0, 0])
    B2 = B[[1, 0, 2]]
    U2, V2 = _aca(B2, 1e-30, max_iter=1)
    assert np.array_equal(V2[0], B2[0] / B2[0, 0])


def test_zero_block_propagates_nan_like_the_reference():
    # pivot 0 -> row / 0 -> NaN, and the NaN stop test never fires (aca.jl:92-97)
    U, V = _aca(np.zeros((4, 3)), 1e-6)
    assert U.shape[1] == 3
    assert np.all(np.isnan(V))


def test_qr_against_lapack():
    A = np.random.default_rng(3).standard_normal((50, 12))
    F, tau = O.qr(A)
    R = np.triu(F[:12])
    Rl = sla.qr(A, mode="r")[0][:12]
    # R is unique up to the signs of its rows
    assert np.allclose(np.abs(R), np.abs(Rl), rtol=1e-12, atol=1e-12)
    # Q (applied to the identity) is orthonormal and reproduces A
    Qb = np.zeros((50, 12), order="F")
    Qb[:12, :12] = np.eye(12)
    O.lib.orc_qr_apply(50, 12, F.ctypes.data_as(O._f64p), O.p64(tau), 12, Qb.ctypes.data_as(O._f64p))
    assert np.allclose(Qb.T @ Qb, np.eye(12), atol=1e-13)
    assert np.allclose(Qb @ R, A, atol=1e-12)


@pytest.mark.parametrize("k,seed", [(1, 0), (5, 1), (24, 2), (40, 3)])
def test_jacobi_svd_against_gesdd(k, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((k, k)) * np.logspace(0, -8, k)[None, :]
    W, sig, Z = O.jacobi_svd(M)
    s_ref = sla.svd(M, compute_uv=False, lapack_driver="gesdd")
    assert np.allclose(sig, s_ref, rtol=1e-12, atol=1e-14 * s_ref[0])
    assert np.all(np.diff(sig) <= 0)
    assert np.allclose(W @ np.diag(sig) @ Z.T, M, atol=1e-13 * s_ref[0])
    assert np.allclose(Z.T @ Z, np.eye(k), atol=1e-13)


def test_truncation_rule_on_built_spectrum():
    rng = np.random.default_rng(4)
    sig = np.array([1.0, 1e-3, 1e-7, 1e-9, 1e-12])
    Qu = np.linalg.qr(rng.standard_normal((30, 5)))[0]
    Qv = np.linalg.qr(rng.standard_normal((25, 5)))[0]
    U = Qu * sig
    V = Qv.T
    # tails: sqrt(sum_{j>=i} sig_j^2) -> [~1, ~1e-3, ~1.005e-7, ~1e-9, 1e-12]
    for eps, r in [(1e-6, 2), (1e-8, 3), (1e-10, 4), (1e-13, 5), (2.0, 0)]:
        U2, V2 = O.svd_recompress(U, V, eps)
        assert U2.shape == (30, r) and V2.shape == (r, 25)
        # Eckart-Young: the Frobenius error equals the discarded tail
        tail = np.sqrt(np.sum(sig[r:] ** 2))
        err = np.linalg.norm(U @ V - U2 @ V2)
        assert abs(err - tail) <= 1e-13 + 1e-9 * tail
        assert tail < eps


def test_recompression_of_aca_factors_matches_lapack_spectrum():
    B = _ref_block()
    U, V = _aca(B, 1e-8)
    s_ref = sla.svd(U @ V, compute_uv=False, lapack_driver="gesdd")
    U2, V2 = O.svd_recompress(U, V, 0.0)  # no truncation: all k kept
    s2 = sla.svd(U2 @ V2, compute_uv=False)
    k = U.shape[1]
    assert np.allclose(s2[:k], s_ref[:k], rtol=1e-9, atol=1e-14 * s_ref[0])


def test_sumsq_lanes_pins():
    # integer-valued squares below 2^53 are summed exactly in any order
    x = np.arange(-700, 1300, dtype=np.float64)
    assert O.sumsq_lanes(x) == float(np.sum(x.astype(np.int64) ** 2))
    assert O.sumsq_lanes(np.zeros(0)) == 0.0
    assert O.sumsq_lanes(np.array([3.0])) == 9.0
    # random data: within n ulp of an extended-precision sum
    r = np.random.default_rng(4).standard_normal(5000)
    exact = float(np.sum(r.astype(np.longdouble) ** 2))
    assert abs(O.sumsq_lanes(r) - exact) <= 5000 * np.finfo(float).eps * exact
    # the fold is the documented order: lane partials then p[l] += p[l+off]
    p = np.zeros(256)
    for i, v in enumerate(r):
        p[i % 256] = p[i % 256] + v * v
    off = 128
    while off:
        p[:off] = p[:off] + p[off:2 * off]
        off //= 2
    assert O.sumsq_lanes(r) == p[0]


def test_dot_lanes64_pins():
    a = np.arange(-300, 500, dtype=np.float64)
    b = np.arange(800, 0, -1, dtype=np.float64)
    assert O.dot_lanes64(a, b) == float(np.dot(a.astype(np.int64), b.astype(np.int64)))
    assert O.dot_lanes64(np.zeros(0), np.zeros(0)) == 0.0
    r = np.random.default_rng(5).standard_normal((2, 3000))
    exact = float(np.sum(r[0].astype(np.longdouble) * r[1].astype(np.longdouble)))
    assert abs(O.dot_lanes64(r[0], r[1]) - exact) <= 3000 * np.finfo(float).eps * np.sum(np.abs(r[0] * r[1]))
