"""Shared test helpers: seeded point sets and the reference-layout adapter."""
import types

import numpy as np


def uniform_points(n, seed):
    """n points uniform in [0,1)^3 as a (3, n) array (Julia column per point)."""
    return np.random.default_rng(seed).random((n, 3)).T.copy()


def layout_ns(H):
    """Product layout (Julia 5 x n orientation) -> object the oracle's mul_layout reads."""
    lay = H.layout(True)
    return types.SimpleNamespace(
        target_index_map=lay["target_index_map"], source_index_map=lay["source_index_map"],
        dense_block_indices=lay["dense_block_indices"].T, U_block_indices=lay["U_block_indices"].T,
        V_block_indices=lay["V_block_indices"].T, dense_blocks=lay["dense_blocks"],
        U_matrices=lay["U_matrices"], V_matrices=lay["V_matrices"])


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def box_instances(count, nmin, nmax, seed):
    """Uneven instances: N log-uniform in [nmin, nmax], points uniform in a box whose
    side along one random axis is a ~ U[1, 8] (the others 1).  Returns a list of
    (3, N) arrays."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))))
        side = np.ones(3)
        side[rng.integers(3)] = rng.uniform(1.0, 8.0)
        out.append((rng.random((n, 3)) * side).T.copy())
    return out
