"""The sensor transactionisation restatement (CPU only): closed forms and hand cases."""
import numpy as np

from oracle import transact as X


def test_balanced_bins_of_distinct_values():
    n, B = 1200, 4
    col = np.random.default_rng(0).permutation(n).astype(np.float64)
    e = X.fit_edges(col, B)
    assert np.allclose(e, [0, 299.75, 599.5, 899.25, 1199])
    counts = np.bincount(X.bin_of(col, e), minlength=B)
    assert counts.tolist() == [300, 300, 300, 300]      # the top value goes to the last bin (clamp)


def test_constant_column_and_ties():
    e = X.fit_edges(np.full(10, 3.0), 5)
    assert e.tolist() == [2.5, 3.5]
    assert X.bin_of(np.array([3.0, -1.0, 9.0]), e).tolist() == [0, 0, 0]
    e = X.fit_edges(np.array([1.0, 1.0, 1.0, 2.0]), 3)   # quantiles 1, 1, 1.0, 2 -> unique [1, 2]
    assert e.tolist() == [1.0, 2.0]


def test_window_medians_by_hand():
    x = np.array([[1.0, 10.0], [3.0, 30.0], [2.0, 20.0], [8.0, 80.0], [5.0, 50.0]])
    off, ids = X.transactions(x, 2, window=2, stride=2)   # windows [0,2), [2,4)
    assert off.tolist() == [0, 2, 4]
    # column 0 edges: quantile(.5) of 1,3,2,8,5 = 3 -> [1, 3, 8]; medians 2 and 5 -> bins 0, 1
    # column 1 edges [10, 30, 80]; medians 20 and 50 -> bins 0, 1
    assert ids.tolist() == [1, 3, 2, 4]
