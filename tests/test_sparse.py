"""Sparse cell list, CPU checks of the oracle's restatement: a hand-made dense
list with known non-empty cells and neighbours, and on random dense lists the
sparse table's candidates equal the dense 27-cell walk's for every non-empty
cell."""
import numpy as np
import pytest

from oracle import oracle as ora


def test_known_answer_one_row():
    # 7 x 1 x 1 cells holding 0, 2, 0, 3, 0, 0, 1 particles
    cs = np.array([0, 0, 2, 2, 5, 5, 5, 6])
    key, start, nn = ora.cl_sparse(cs, (7, 1, 1))
    assert key.tolist() == [1, 3, 6] and start.tolist() == [0, 2, 5, 6]
    # along x (dy = dz = 0: k = 12, 13, 14): left, self, right
    assert nn[:, 12:15].tolist() == [[-1, 0, -1], [-1, 1, -1], [-1, 2, -1]]
    assert np.all(np.delete(nn, [12, 13, 14], axis=1) == -1)  # one row of cells: no y/z neighbours


@pytest.mark.parametrize("box_nn", [1, 2, 3])
def test_random_dense_lists_walk_equal(box_nn):
    rng = np.random.default_rng(1)
    B = range(-box_nn, box_nn + 1)
    for divc in ((5, 4, 3), (3, 3, 3), (8, 2, 5), (9, 7, 8)):
        n_cells = int(np.prod(divc))
        counts = rng.integers(0, 3, n_cells) * (rng.random(n_cells) < 0.4)
        cs = np.concatenate([[0], np.cumsum(counts)])
        key, start, nn = ora.cl_sparse(cs, divc, box_nn)
        assert np.array_equal(key, np.flatnonzero(counts))
        assert nn.shape == (len(key), (2 * box_nn + 1) ** 3)
        d0, d1, d2 = divc
        for m, c in enumerate(key):
            cx, cy, cz = c % d0, (c // d0) % d1, c // (d0 * d1)
            dense = []
            for dz in B:
                for dy in B:
                    for dx in B:
                        x, y, z = cx + dx, cy + dy, cz + dz
                        if 0 <= x < d0 and 0 <= y < d1 and 0 <= z < d2:
                            q = x + d0 * (y + d1 * z)
                            dense.extend(range(cs[q], cs[q + 1]))
            sparse = [i for e in nn[m] if e >= 0 for i in range(start[e], start[e + 1])]
            assert sparse == dense
