"""Pins for the oracle's post-processing filters.

The reference declares speckle removal, gap interpolation, adaptive-mean and
median filtering (src/Elas.h:31-35) but has no code for them, so there is no
reference output: each definition is pinned here by a library routine that
computes the same thing independently (scipy connected components and median
filter), by hand-computed cases, or by invariants."""
import numpy as np
import pytest
from scipy import ndimage
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components


def _speckle_by_components(D, sim, max_size):
    """Speckle removal through scipy's connected components of the
    4-neighbour graph with edges where both pixels are valid and
    |d_a - d_b| <= sim (float32 arithmetic, as the filter)."""
    H, W = D.shape
    idx = np.arange(H * W).reshape(H, W)
    valid = D >= 0
    rows, cols = [], []
    f32 = np.float32
    for a, b, da, db, va, vb in ((idx[:, :-1], idx[:, 1:], D[:, :-1], D[:, 1:], valid[:, :-1], valid[:, 1:]),
                                 (idx[:-1, :], idx[1:, :], D[:-1, :], D[1:, :], valid[:-1, :], valid[1:, :])):
        e = va & vb & (np.abs((da - db).astype(f32)) <= f32(sim))
        rows.append(a[e])
        cols.append(b[e])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    g = coo_matrix((np.ones(len(r)), (r, c)), shape=(H * W, H * W))
    _, lab = connected_components(g, directed=False)
    sizes = np.bincount(lab[valid.ravel()], minlength=H * W)
    out = D.copy().ravel()
    small = valid.ravel() & (sizes[lab] < max_size)
    out[small] = -1
    return out.reshape(H, W)


def _random_map(rng, H, W, levels, p_invalid, smooth=False):
    D = rng.integers(0, levels, (H, W)).astype(np.float32)
    if smooth:
        D = np.cumsum(rng.integers(-1, 2, (H, W)), axis=1).astype(np.float32) * 0.5 + 20
        D += rng.random((H, W)).astype(np.float32) * 0.25
    D[rng.random((H, W)) < p_invalid] = -1
    return D.astype(np.float32)


@pytest.mark.parametrize("seed,sim,max_size,levels,smooth", [
    (0, 0.0, 5, 3, False), (1, 1.0, 12, 6, False), (2, 1.0, 40, 4, True), (3, 2.5, 200, 30, True),
    (4, 0.5, 3, 2, False)])
def test_speckle_equals_connected_components(oracle, seed, sim, max_size, levels, smooth):
    rng = np.random.default_rng(seed)
    D = _random_map(rng, 37, 53, levels, 0.2, smooth)
    got = oracle.speckle(D, sim, max_size)
    assert np.array_equal(got, _speckle_by_components(D, sim, max_size))


def test_gap_interpolation_hand_cases(oracle):
    def row(vals, gap):
        return oracle.gap_interpolation(np.array([vals], np.float32), gap)[0].tolist()
    assert row([2, -1, -1, 4], 5) == [2, 3, 3, 4]                 # |2-4| < 3: mean
    assert row([2, -1, 9], 5) == [2, 2, 9]                        # else the smaller
    assert row([2, -1, -1, 4.5], 5) == [2, 3.25, 3.25, 4.5]
    assert row([2, -1, 5], 5) == [2, 2, 5]                        # |2-5| = 3 is not < 3
    assert row([-1, 3, -1], 5) == [-1, 3, -1]                     # runs touching the border stay
    assert row([1] + [-1] * 6 + [1], 5) == [1] + [-1] * 6 + [1]   # run longer than the gap
    assert row([1] + [-1] * 6 + [1], 6) == [1] * 8
    # rows first, then columns of the row-filled map
    D = np.array([[1, -1], [-1, -1], [3, -1]], np.float32)
    assert oracle.gap_interpolation(D, 5).tolist() == [[1, -1], [2, -1], [3, -1]]
    D = np.array([[4, -1, 4], [-1, -1, -1], [8, -1, 8]], np.float32)
    # row 1 stays invalid after the row pass; the column pass fills it with min(4, 8)
    assert oracle.gap_interpolation(D, 5).tolist() == [[4, 4, 4], [4, 4, 4], [8, 8, 8]]


def test_gap_interpolation_invariants(oracle):
    rng = np.random.default_rng(7)
    D = _random_map(rng, 40, 60, 50, 0.35)
    for gap in (1, 3, 10, 1000):
        got = oracle.gap_interpolation(D, gap)
        v = D >= 0
        assert np.array_equal(got[v], D[v])                       # valid pixels never change
        assert (got >= -1).all()
        filled = (~v) & (got >= 0)
        # every filled value is one of the valid values (or a mean of two) in its row/column
        assert filled.sum() > 0 if gap >= 3 else True
        lo, hi = D[v].min(), D[v].max()
        assert (got[filled] >= lo).all() and (got[filled] <= hi).all()


def test_adaptive_mean_invariants(oracle):
    D = np.full((20, 30), 17.0, np.float32)
    D[5, 5] = -1
    got = oracle.adaptive_mean(D)
    assert np.array_equal(got, D)                                 # constant: unchanged, invalid stays
    step = np.full((20, 30), 10.0, np.float32)
    step[:, 15:] = 30.0                                           # jump >= 4: no weight across it
    got = oracle.adaptive_mean(step)
    assert np.array_equal(got, step)
    u = np.arange(40, dtype=np.float32)
    ramp = np.tile(u * 0.5 + 3, (25, 1)).astype(np.float32)       # symmetric weights: interior unchanged
    got = oracle.adaptive_mean(ramp)
    assert np.allclose(got[:, 3:-3], ramp[:, 3:-3], atol=1e-5)
    rng = np.random.default_rng(3)
    R = _random_map(rng, 30, 30, 12, 0.2)
    got = oracle.adaptive_mean(R)
    assert np.array_equal(got < 0, R < 0)
    assert got[R >= 0].min() >= R[R >= 0].min() and got.max() <= R.max()


def test_adaptive_mean_hand_cases(oracle):
    """Hand-computed values that fix the window radius (3), the weight
    max(0, 4 - |d_k - d_c|) and the order (rows, then columns)."""
    # one row: the centre (index 4, d = 0) sees index 1 (d = 2, w = 2), itself
    # (w = 4) and index 7 (d = 3, w = 1); the 60s weigh nothing and indices 0
    # and 8 (d = 3) lie outside the radius: (2*2 + 0*4 + 3*1) / (2 + 4 + 1) = 1
    row = np.array([[3, 2, 60, 60, 0, 60, 60, 3, 3]], np.float32)
    got = oracle.adaptive_mean(row)
    assert got[0, 4] == 1.0
    # index 7 (d = 3): index 4 (d = 0, w = 1), itself (4), index 8 (d = 3, w = 4);
    # index 10 does not exist: (0*1 + 3*4 + 3*4) / 9
    assert abs(got[0, 7] - 24.0 / 9.0) < 1e-6
    # the same as a column
    col = oracle.adaptive_mean(row.T.copy())
    assert np.array_equal(col[:, 0], got[0])
    # 2 x 2: rows first -> [[3/7, 4/7], [2, 6]]; then column 0 (3/7 and 2,
    # |delta| = 11/7, w = 17/7): top (4*3/7 + 17/7*2) / (4 + 17/7) = 46/45,
    # bottom (4*2 + 17/7*3/7) / (45/7) = 443/315; column 1 differs by 38/7 > 4
    D = oracle.adaptive_mean(np.array([[0, 1], [2, 6]], np.float32))
    exp = np.array([[46 / 45, 4 / 7], [443 / 315, 6]])
    assert np.allclose(D, exp, rtol=0, atol=2e-6), D
    # invalid neighbours are skipped, not counted as d = -1
    D = oracle.adaptive_mean(np.array([[1, -1, 2]], np.float32))
    # index 0: itself (4) and index 2 (|1| -> w = 3): (4 + 6) / 7
    assert abs(D[0, 0] - 10 / 7) < 1e-6 and D[0, 1] == -1


def test_median_equals_scipy_on_valid_interior(oracle):
    rng = np.random.default_rng(5)
    D = rng.integers(0, 100, (40, 50)).astype(np.float32)
    got = oracle.median(D)
    ref = ndimage.median_filter(ndimage.median_filter(D, size=(1, 7)), size=(7, 1))
    assert np.array_equal(got[3:-3, 3:-3], ref[3:-3, 3:-3])


def test_median_hand_cases(oracle):
    D = oracle.median(np.array([[1, -1, 5, 9, 2, -1, 7]], np.float32))
    # centre 5 (index 2): window valid {1, 5, 9, 2}   -> sorted 1 2 5 9, lower middle 2
    # index 3 (9): window {1, 5, 9, 2, 7}            -> 5
    assert D[0, 2] == 2 and D[0, 3] == 5 and D[0, 1] == -1 and D[0, 5] == -1
