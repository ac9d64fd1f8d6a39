"""The restatement's OpenMP loops (cell and face loops of blocks above 4,096
cells, and the cells of each LU-SGS hyperplane) give the same bits at any
thread count: every parallel iteration writes only its own cell or face and
every sum keeps its serial order.  A Spalart-Allmaras vane passage whose blocks
are all above the threshold runs its two-grid start-up and cycle on one and on
eight threads."""
import numpy as np

from fv3d import cases
import oracle as O


def test_thread_count_does_not_change_results():
    case = cases.vane_passage(ni=24, nk_blocks=(8, 8, 16, 8, 8), levels=0, coarse_iter=2, max_iter=4)
    assert min((b.im - 1) * (b.jm - 1) * (b.km - 1) for b in case.blocks) > 4096
    runs = []
    for threads in (1, 8):
        O.set_threads(threads)
        o = O.Oracle(case)
        h = o.iterate(4)
        runs.append((h, [o.get(b, "cv").copy() for b in range(1, len(case.blocks) + 1)],
                     [o.get(b, "ev").copy() for b in range(1, len(case.blocks) + 1)]))
        del o
    O.set_threads(0)
    (h1, cv1, ev1), (h8, cv8, ev8) = runs
    assert np.array_equal(h1, h8)
    for a, b in zip(cv1 + ev1, cv8 + ev8):
        assert np.array_equal(a, b)
