"""The oracle's restatement of compressed2D_inserter (orc_assemble) against
T-mtl/compressed2D_inserter_test.cpp: a chain of inserters on an 8 x 6 matrix
with update_store and update_plus, each starting from the previous matrix."""
import numpy as np

import oracle_lib as o

INSERTER_STEPS = [
    # (updater, [(row, col, value), ...]) in the test's insertion order
    (0, [(2, 2, 6), (7, 2, 17)]),
    (0, [(0, 3, 31), (3, 3, 33), (6, 0, 34), (6, 0, 35), (4, 4, 36), (4, 4, 37)]),
    (1, [(2, 2, 21), (2, 4, 22), (6, 1, 23), (7, 2, 24), (7, 2, 2), (4, 2, 25), (2, 5, 26), (0, 2, 27),
         (3, 1, 28), (4, 2, 29)]),
    (1, [(2, 2, 1)]),
]


def run_steps(assemble):
    A, after = None, []
    for upd, ins in INSERTER_STEPS:
        r, c, v = (np.array([t[k] for t in ins]) for k in range(3))
        A = assemble(r, c, v.astype(float), upd, A)
        after.append(A)
    return after


def entry(A, r, c):
    for j in range(A.starts[r], A.starts[r + 1]):
        if A.indices[j] == c:
            return A.data[j]
    return 0.0


def test_inserter_known_answers():
    after = run_steps(lambda r, c, v, u, A: o.assemble(8, r, c, v, u, A))
    B = after[1]
    assert entry(B, 0, 3) == 31 and entry(B, 3, 3) == 33 and entry(B, 6, 0) == 35 and entry(B, 4, 4) == 37
    C = after[2]
    assert entry(C, 0, 2) == 27 and entry(C, 2, 2) == 27 and entry(C, 4, 2) == 54 and entry(C, 7, 2) == 43
    assert entry(after[3], 2, 2) == 28
    # rows sorted, no duplicates
    for r in range(8):
        cols = after[3].indices[after[3].starts[r]:after[3].starts[r + 1]]
        assert list(cols) == sorted(set(cols))


def test_inserter_minus_and_times():
    # update_minus stores -v for a new entry and subtracts afterwards; update_times multiplies
    A = o.assemble(2, [0, 0, 1], [1, 1, 0], [2.0, 3.0, 4.0], updater=2)
    assert entry(A, 0, 1) == -5.0 and entry(A, 1, 0) == -4.0
    B = o.assemble(2, [0, 1, 1], [1, 0, 0], [3.0, 0.5, 3.0], updater=3, existing=A)
    assert entry(B, 0, 1) == -15.0 and entry(B, 1, 0) == -6.0
