"""Host Delaunay step: for integer points without duplicates the library
builds Triangle's divide-and-conquer vertex order itself (two counting sorts
and alternating median cuts) and runs Triangle's own merge from there; other
sets go through Triangle unchanged.  The meshes, triangle numbering included,
must be those of Triangle on the caller's order.

Checked against the oracle's own build of the reference's triangle.cpp, which
always triangulates the points in the caller's (row-major) order, as the
reference does (src/Elas.cu:278-328).  No device needed.
"""
import numpy as np
import pytest


def _lattice(rng, cw, ch, step, keep):
    cu, cv = np.meshgrid(np.arange(cw), np.arange(ch))
    m = rng.random((ch, cw)) < keep
    return np.stack([cu[m] * step, cv[m] * step], 1).astype(np.float32)


def test_golden_mesh(pkg):
    z = np.load(__file__.rsplit("/", 1)[0] + "/golden/small_pair.npz")
    assert np.array_equal(pkg.triangulate(z["sp0"]), z["tri0"])


@pytest.mark.parametrize("seed", range(6))
def test_lattice_sets_equal_reference_order(pkg, oracle, seed):
    """Left-image support points: lattice cells in row order (highly cocircular)."""
    rng = np.random.default_rng(seed)
    pts = _lattice(rng, 40 + 17 * seed, 20 + 9 * seed, 5, 0.3 + 0.1 * seed)
    assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts))


@pytest.mark.parametrize("seed", range(6))
def test_shifted_sets_with_duplicates_equal_reference_order(pkg, oracle, seed):
    """Right-image support points (u - d, v): duplicates occur where two cells of
    a row differ in d by their u distance; the library then runs Triangle's own
    sort and duplicate removal on the caller's order, since which duplicate
    Triangle drops depends on that order."""
    rng = np.random.default_rng(100 + seed)
    pts = _lattice(rng, 60, 30, 5, 0.6)
    d = rng.integers(0, 12, len(pts)).astype(np.float32)
    pts[:, 0] -= d
    if seed % 2 == 0:  # force duplicates
        pts[1::7] = pts[0:-1:7][: len(pts[1::7])]
    assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts))


def test_all_points_equal_is_too_few_points(pkg):
    """Three or more copies of one point leave a single vertex after duplicate
    removal, where Triangle's divconqrecurse() recurses until the stack
    overflows (the oracle's Triangle is not called here for that reason): the
    library reports too few points instead."""
    for n in (3, 5):
        with pytest.raises(pkg.ElasError) as e:
            pkg.triangulate(np.full((n, 2), 7.0, np.float32))
        assert e.value.code == pkg.ELAS_ERR_TOO_FEW_POINTS


def test_two_distinct_points_after_duplicates_give_an_empty_mesh(pkg, oracle):
    pts = np.array([[0, 0], [0, 0], [5, 5], [5, 5], [0, 0]], np.float32)
    assert len(oracle.triangulate(pts)) == 0
    assert len(pkg.triangulate(pts)) == 0


def test_order_independence_without_duplicates(pkg, oracle):
    """The property the sorted hand-over relies on: for distinct points,
    Triangle's mesh (corners numbered by input position) is the same for any
    input order."""
    rng = np.random.default_rng(7)
    pts = _lattice(rng, 70, 40, 5, 0.5)
    ref = oracle.triangulate(pts)
    for rep in range(4):
        perm = rng.permutation(len(pts))
        got = pkg.triangulate(pts[perm])
        assert np.array_equal(perm[got], ref)


def test_non_integer_points_keep_caller_order(pkg, oracle):
    rng = np.random.default_rng(3)
    pts = (rng.random((500, 2)) * [300, 200]).astype(np.float32)
    assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts))


@pytest.mark.parametrize("k", [0, 3, 7, 8, 15, 100, 200, 202])
def test_one_non_integer_coordinate_anywhere(pkg, oracle, k):
    """The integer keys are converted eight points at a time, the last n % 8
    one by one: a single half-integer coordinate in either part must send the
    set to Triangle's own sort (taken as an integer it would tie with, and be
    ordered among, the points of its truncated column)."""
    rng = np.random.default_rng(40 + k)
    flat = rng.choice(30 * 30, 203, replace=False)
    pts = np.stack([flat % 30, flat // 30], 1).astype(np.float32)
    pts[k, k % 2] += 0.5
    assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts))


def test_pipeline_meshes_of_c0(pkg, oracle, pairs):
    left, right, _, dmax = pairs("450x375_d64")
    ref = oracle.process(left, right, oracle.params(disp_max=dmax))
    assert np.array_equal(pkg.triangulate(ref["sp0"]), ref["tri0"])
    assert np.array_equal(pkg.triangulate(ref["sp1"]), ref["tri1"])


@pytest.mark.parametrize("n", [3, 4, 5, 6, 7, 8, 9, 12, 13, 31, 64, 65, 257, 1000, 4001])
def test_random_integer_sets_every_size(pkg, oracle, n):
    """Sizes that exercise every leaf shape (2 and 3 points) and odd halves of
    the alternating cuts; coordinates from a small range, so many points share
    an x or a y and the cuts' tie order (other axis) decides."""
    rng = np.random.default_rng(n)
    side = max(4, int(np.sqrt(n) * 2))
    flat = rng.choice(side * side, n, replace=False)
    pts = np.stack([flat % side, flat // side], 1).astype(np.float32) * 3 - 7
    assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts))


@pytest.mark.parametrize("n", [65536, 65537, 70001])
def test_integer_sets_at_and_past_16_bit_ranks(pkg, oracle, n):
    """The vertex order packs (x rank, y rank) into 32-bit entries up to
    65536 points and into 64-bit entries beyond: both sides of the switch
    give Triangle's mesh."""
    rng = np.random.default_rng(n)
    side = 400
    flat = rng.choice(side * side, n, replace=False)
    pts = np.stack([flat % side, flat // side], 1).astype(np.float32) * 2
    assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts))


def test_full_lattice_cocircular(pkg, oracle):
    """Every 4 neighbours of a full lattice are cocircular: the mesh depends
    on the exact merge order, which the vertex order fixes."""
    cu, cv = np.meshgrid(np.arange(37), np.arange(23))
    pts = np.stack([cu.ravel(), cv.ravel()], 1).astype(np.float32) * 5
    assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts))


def test_collinear_points(pkg, oracle):
    for pts in (np.stack([np.arange(9) * 5, np.zeros(9)], 1), np.stack([np.zeros(7), np.arange(7) * 2], 1),
                np.stack([np.arange(6), np.arange(6)], 1)):
        pts = pts.astype(np.float32)
        assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts))


def test_pipeline_meshes_of_c1(pkg, oracle, pairs):
    """The 1242x375 pair's support points (about 1,400 per image)."""
    left, right, _, dmax = pairs("1242x375_d128")
    ref = oracle.process(left, right, oracle.params(disp_max=dmax))
    assert np.array_equal(pkg.triangulate(ref["sp0"]), ref["tri0"])
    assert np.array_equal(pkg.triangulate(ref["sp1"]), ref["tri1"])


@pytest.mark.parametrize("seed", range(4))
def test_heavy_duplicates(pkg, oracle, seed):
    """Every point repeated up to 4 times in shuffled order, integer and
    non-integer: Triangle's own sort decides which copy survives (the others
    are dropped), and only the cuts of the survivors are built outside."""
    rng = np.random.default_rng(500 + seed)
    base = rng.integers(0, 60, (300, 2)).astype(np.float32)
    if seed % 2:
        base = base * 0.37 + 0.11
    pts = np.repeat(base, rng.integers(1, 5, len(base)), axis=0)
    pts = pts[rng.permutation(len(pts))]
    assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts))


def test_too_few_points(pkg):
    with pytest.raises(pkg.ElasError) as e:
        pkg.triangulate(np.zeros((2, 2), np.float32))
    assert e.value.code == pkg.ELAS_ERR_TOO_FEW_POINTS


@pytest.fixture
def split(pkg):
    """Forces the split depth of the merge over the host threads; restores
    the automatic setting afterwards."""
    yield pkg.set_tri_split
    pkg.set_tri_split(-1)


def _split_sets(oracle_rng_seed=0):
    rng = np.random.default_rng(oracle_rng_seed)
    sets = [("lattice", _lattice(rng, 61, 37, 5, 0.55))]
    for n in (4, 7, 13, 64, 65, 257, 1000, 3001):
        side = max(4, int(np.sqrt(n) * 2))
        flat = rng.choice(side * side, n, replace=False)
        sets.append(("random%d" % n, np.stack([flat % side, flat // side], 1).astype(np.float32) * 3 - 7))
    cu, cv = np.meshgrid(np.arange(41), np.arange(29))
    sets.append(("full", np.stack([cu.ravel(), cv.ravel()], 1).astype(np.float32) * 5))
    sets.append(("collinear", np.stack([np.arange(40) * 5, np.zeros(40)], 1).astype(np.float32)))
    base = rng.integers(0, 60, (400, 2)).astype(np.float32)
    dup = np.repeat(base, rng.integers(1, 4, len(base)), axis=0)
    sets.append(("duplicates", dup[rng.permutation(len(dup))]))
    sets.append(("nonint", (rng.random((900, 2)) * [300, 200]).astype(np.float32)))
    return sets


@pytest.mark.parametrize("depth", [0, 1, 2, 3, 4, 6])
def test_split_merge_equals_serial_triangle(pkg, oracle, split, depth):
    """The merge cut into up to 2^depth concurrent subtrees (each with its own
    triangle pool, read out in post-order) gives Triangle's serial triangle
    list on every kind of set: cocircular lattices, collinear points, sets
    with duplicates and non-integer coordinates.  Small subtrees are forced
    (the setting's lower bound on subtree size applies only to the automatic
    depth), so leaves of 2 and 3 points meet the merge too."""
    split(depth)
    for name, pts in _split_sets(depth):
        assert pkg.tri_split_depth(len(pts)) == (min(depth, int(np.log2(len(pts) / 4))) if len(pts) >= 8 else 0)
        assert np.array_equal(pkg.triangulate(pts), oracle.triangulate(pts)), (name, depth)


def test_split_on_pipeline_points(pkg, oracle, pairs, split):
    left, right, _, dmax = pairs("1242x375_d128")
    ref = oracle.process(left, right, oracle.params(disp_max=dmax))
    for depth in (1, 2, 3, 5):
        split(depth)
        assert np.array_equal(pkg.triangulate(ref["sp0"]), ref["tri0"]), depth
        assert np.array_equal(pkg.triangulate(ref["sp1"]), ref["tri1"]), depth
    split(-1)
    assert pkg.tri_split_depth(len(ref["sp0"])) == 3        # 1,427 points: 8 subtrees of >= 128
    assert pkg.tri_split_depth(22418) == 3                   # capped at 8 subtrees
