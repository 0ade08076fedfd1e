"""Pins for the oracle's cluster tree (tree.jl:45-187) and block tree
(block.jl:31-155): hand-computed cases, invariants and brute force.

The reference's own tests restated here:
  test_tree.jl:6-15          1000 random planar points, leaf 10: max <= 10, min >= 3
  test_block.jl:19-97        root split, childless labels, admissibility at
                             eta = 0.5, partition of unity (exact coverage)
  test_hmatrix_vector.jl:33-35  dims=3 index expansion is a permutation of 1..3N
"""
import numpy as np
import pytest

import oracle_lib as O


def test_collinear_hand_computed_split():
    # x = 0, 1, 2, 3, 10 on a line: centroid 3.2 splits {0,1,2,3} | {10};
    # the left child (4 points > leaf 2) splits at 1.5 into {0,1} | {2,3}.
    pts = np.array([[10, 0, 0], [0, 0, 0], [3, 0, 0], [1, 0, 0], [2, 0, 0]], float)
    t = O.Tree(pts, leaf=2)
    # index_map: stable partition keeps input order inside each side
    assert t.index_map.tolist() == [2, 4, 3, 5, 1]
    assert t.start.tolist() == [1, 1, 1, 3, 5]
    assert t.end.tolist() == [6, 5, 3, 5, 6]
    assert t.is_leaf.tolist() == [0, 0, 1, 1, 1]
    assert t.left.tolist() == [1, 2, -1, -1, -1]
    assert t.right.tolist() == [4, 3, -1, -1, -1]
    # root: centroid 16/5 = 3.2, radius = max(|3.2-10|, |3.2-0|) = 6.8
    assert t.center[0].tolist() == [16 / 5, 0.0, 0.0]
    assert t.radius[0] == abs(16 / 5 - 10)
    # node {0,1,3,2} (in index order 1,3,4,5 -> x = 0,3,1,2): centroid 1.5, radius 1.5
    assert t.center[1, 0] == 1.5 and t.radius[1] == 1.5
    # leaf {x=0,1}: centroid 0.5, radius 0.5; leaf {x=3,2}: centroid 2.5, radius 0.5
    assert t.center[2, 0] == 0.5 and t.radius[2] == 0.5
    assert t.center[3, 0] == 2.5 and t.radius[3] == 0.5


def test_ties_go_right_and_first_longest_axis():
    # Equal box lengths in x and y: argmax picks x (first maximum).  A point
    # exactly at the centroid goes to the right child (strict <).
    pts = np.array([[0, 0, 0], [2, 2, 0], [1, 1, 0]], float)
    t = O.Tree(pts, leaf=2)
    assert t.index_map.tolist() == [1, 2, 3]
    # centroid x = 1: only point 1 (x=0) is strictly left
    root_left = t.left[0]
    assert t.end[root_left] - t.start[root_left] == 1
    assert t.index_map[0] == 1
    # right child holds points 2 and 3 in their original relative order
    assert t.index_map[1:].tolist() == [2, 3]


def _check_tree_invariants(t, pts, leaf):
    n = len(pts)
    assert sorted(t.index_map.tolist()) == list(range(1, n + 1))
    for k in range(len(t.start)):
        s, e = t.start[k], t.end[k]
        ids = t.index_map[s - 1:e - 1] - 1
        P = pts[ids]
        lo, hi = P.min(0), P.max(0)
        c = t.center[k]
        # centroid inside the box, radius reaches the farther box corner only
        assert np.all(c >= lo - 1e-12) and np.all(c <= hi + 1e-12)
        far = max(np.linalg.norm(c - hi), np.linalg.norm(c - lo))
        assert abs(t.radius[k] - far) <= 1e-15 * max(1.0, far)
        assert abs(np.linalg.norm(c - P.mean(0))) <= 1e-12
        if t.is_leaf[k]:
            assert e - s <= leaf and t.left[k] == -1
        else:
            assert e - s > leaf
            l, r = t.left[k], t.right[k]
            assert t.start[l] == s and t.end[l] == t.start[r] and t.end[r] == e
            axis = int(np.argmax(hi - lo))
            L = pts[t.index_map[t.start[l] - 1:t.end[l] - 1] - 1, axis]
            R = pts[t.index_map[t.start[r] - 1:t.end[r] - 1] - 1, axis]
            assert np.all(L < c[axis]) and np.all(R >= c[axis])


@pytest.mark.parametrize("n,leaf,seed", [(300, 8, 1), (1000, 10, 2), (2048, 32, 3)])
def test_tree_invariants_brute_force(n, leaf, seed):
    pts = np.random.default_rng(seed).random((n, 3))
    t = O.Tree(pts, leaf)
    _check_tree_invariants(t, pts, leaf)


def test_partition_is_stable():
    # one split of an input ordered by decreasing x: both sides keep the
    # input order (increasing ids) and the left side holds the larger ids
    pts = np.stack([np.linspace(0, 1, 40), np.zeros(40), np.zeros(40)], 1)[::-1].copy()
    t = O.Tree(pts, 39)
    left = t.index_map[t.start[t.left[0]] - 1:t.end[t.left[0]] - 1]
    right = t.index_map[t.start[t.right[0]] - 1:t.end[t.right[0]] - 1]
    assert np.all(np.diff(left) > 0) and np.all(np.diff(right) > 0)
    assert left.min() > right.max()  # input was descending in x


def test_reference_tree_test_planar_points():
    # test_tree.jl:6-15 with this suite's seeded points
    pts = np.random.default_rng(10).random((1000, 3))
    pts[:, 2] = 0
    info = O.Tree(pts, 10).info()
    assert info["max_points"] <= 10
    assert info["min_points"] >= 3


def test_dims_expansion():
    pts = np.random.default_rng(5).random((200, 3))
    t1 = O.Tree(pts, 16)
    t3 = O.Tree(pts, 16, dims=3)
    assert sorted(t3.index_map.tolist()) == list(range(1, 601))
    # [3p-2, 3p-1, 3p] per point p, in the dims=1 order (tree.jl:59-64)
    exp = np.stack([3 * t1.index_map - 2, 3 * t1.index_map - 1, 3 * t1.index_map], 1).ravel()
    assert np.array_equal(t3.index_map, exp)
    assert np.array_equal(t3.start, 3 * (t1.start - 1) + 1)
    assert np.array_equal(t3.end, 3 * (t1.end - t1.start) + t3.start)


def test_tree_errors():
    pts = np.random.default_rng(6).random((10, 3))
    with pytest.raises(ValueError, match="code 1"):
        O.Tree(pts, 4, dims=0)
    dup = np.zeros((9, 3))
    with pytest.raises(ValueError, match="code 2"):
        O.Tree(dup, 4)


def test_norm3_closed_forms_and_scaling():
    f = lambda v: O.lib.orc_norm3(np.ascontiguousarray(v, float).ctypes.data_as(O._f64p))  # noqa: E731
    assert f([3.0, 4.0, 12.0]) == 13.0
    assert f([0.0, 0.0, 0.0]) == 0.0
    assert f([-2.0, 0.0, 0.0]) == 2.0
    # maxabs^2 underflows: the scaled branch keeps full precision
    assert f([3e-200, 4e-200, 0.0]) == pytest.approx(5e-200, rel=1e-15)
    assert f([3e200, 4e200, 0.0]) == pytest.approx(5e200, rel=1e-15)
    assert f([np.inf, 1.0, 0.0]) == np.inf


# ---- block tree --------------------------------------------------------------------
def _blocks(seed=10, n=32):
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 3))
    pts[:, 2] = 0
    S = O.Tree(pts, 4)
    T = O.Tree(pts, 8)
    return pts, T, S


def test_block_tree_labels_and_admissibility():
    pts, T, S = _blocks()
    for eta in (1.5, 0.5):
        b = O.BlockTree(T, S, eta)
        assert b.left[0] >= 0 and b.right[0] >= 0  # test_block.jl:19-20
        for k in range(len(b.tnode)):
            if b.adm[k] or b.dense[k]:
                assert b.left[k] == -1 and b.right[k] == -1  # test_block.jl:23-31
            if b.adm[k]:  # test_block.jl:34-45
                d = np.linalg.norm(T.center[b.tnode[k]] - S.center[b.snode[k]])
                assert d > eta * (T.radius[b.tnode[k]] + S.radius[b.snode[k]])
            if b.dense[k]:
                assert T.is_leaf[b.tnode[k]] and S.is_leaf[b.snode[k]]


@pytest.mark.parametrize("eta,merge", [(1.5, False), (1.5, True), (0.5, True), (2.0, True)])
def test_partition_of_unity_brute_force(eta, merge):
    """Every (i, j) is covered by exactly one listed block (test_block.jl:71-97)."""
    pts = np.random.default_rng(11).random((400, 3))
    T = O.Tree(pts, 16)
    S = O.Tree(pts, 8)
    b = O.BlockTree(T, S, eta, merge=merge)
    d, a = b.traverse()
    cover = np.zeros((400, 400), np.int32)
    for t, s in np.concatenate([d, a]):
        rows = T.index_map[T.start[t] - 1:T.end[t] - 1] - 1
        cols = S.index_map[S.start[s] - 1:S.end[s] - 1] - 1
        cover[np.ix_(rows, cols)] += 1
    assert np.all(cover == 1)
    # K x equals the sum of the block products (the reference's numeric form)
    K = np.random.default_rng(12).random((400, 400))
    x = np.random.default_rng(13).random(400)
    y = np.zeros(400)
    for t, s in np.concatenate([d, a]):
        rows = T.index_map[T.start[t] - 1:T.end[t] - 1] - 1
        cols = S.index_map[S.start[s] - 1:S.end[s] - 1] - 1
        y[rows] += K[np.ix_(rows, cols)] @ x[cols]
    assert np.allclose(K @ x, y, atol=1e-10)


def test_merge_rule_and_dfs_order():
    pts = np.random.default_rng(14).random((600, 3))
    T = O.Tree(pts, 16)
    b0 = O.BlockTree(T, T, 2.0)
    b = O.BlockTree(T, T, 2.0, merge=True)
    for k in range(len(b.tnode)):
        l, r = b.left[k], b.right[k]
        if l >= 0:  # a parent is dense iff both children are (block.jl:98-112)
            assert b.dense[k] == (b.dense[l] and b.dense[r])
        else:
            assert b.dense[k] == b0.dense[k]
    d, a = b.traverse()
    # listed dense nodes never have a dense ancestor (DFS stops at them)
    parent = {}
    for k in range(len(b.tnode)):
        for c in (b.left[k], b.right[k]):
            if c >= 0:
                parent[c] = k
    listed = set()
    for k in range(len(b.tnode)):
        if b.dense[k] or b.adm[k]:
            p, ok = parent.get(k), True
            while p is not None:
                if b.dense[p] or b.adm[p]:
                    ok = False
                p = parent.get(p)
            if ok:
                listed.add(k)
    # DFS pre-order, left before right, not descending below labelled nodes
    dd, aa = [], []

    def walk(k):
        if b.adm[k]:
            aa.append((b.tnode[k], b.snode[k]))
        elif b.dense[k]:
            dd.append((b.tnode[k], b.snode[k]))
        else:
            walk(b.left[k])
            walk(b.right[k])

    walk(0)
    assert listed == {k for k in range(len(b.tnode)) if (b.dense[k] or b.adm[k]) and
                      all(not (b.dense[p] or b.adm[p]) for p in _ancestors(parent, k))}
    assert [tuple(x) for x in d] == dd
    assert [tuple(x) for x in a] == aa


def _ancestors(parent, k):
    out = []
    p = parent.get(k)
    while p is not None:
        out.append(p)
        p = parent.get(p)
    return out
