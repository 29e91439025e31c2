"""Writes fw_golden.npz: all-pairs shortest paths of small random undirected
graphs computed by SciPy's floyd_warshall (scipy.sparse.csgraph), used to pin
the oracle's and the device's Floyd-Warshall.

Run from the repository root:  python tests/golden/make_fw_golden.py
"""
import os

import numpy as np
from scipy.sparse.csgraph import floyd_warshall

HERE = os.path.dirname(os.path.abspath(__file__))


def random_graph(rng, n, degree, integer):
    W = np.full((n, n), np.inf)
    np.fill_diagonal(W, 0.0)
    m = int(n * degree / 2)
    for _ in range(m):
        i, j = rng.integers(0, n, size=2)
        if i == j:
            continue
        w = float(rng.integers(1, 100)) if integer else float(rng.uniform(0.01, 3.0))
        W[i, j] = W[j, i] = min(W[i, j], w)
    return W


def main():
    rng = np.random.default_rng(20261203)
    cases = {}
    for name, n, deg, integer in [("int16", 16, 3.0, True), ("int64", 64, 4.0, True), ("int200", 200, 3.0, True),
                                  ("flt64", 64, 5.0, False), ("sparse130", 130, 1.2, True)]:
        W = random_graph(rng, n, deg, integer)
        Wz = np.where(np.isinf(W), 0.0, W)  # csgraph dense input: 0 = no edge off the diagonal
        Dm = floyd_warshall(Wz, directed=False)
        cases[name + "_W"] = W.astype(np.float64)
        cases[name + "_D"] = Dm.astype(np.float64)
    np.savez_compressed(os.path.join(HERE, "fw_golden.npz"), **cases)
    print("wrote", sorted(cases))


if __name__ == "__main__":
    main()
