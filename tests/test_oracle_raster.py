"""Pins of the oracle's plane rasteriser (eo_raster) that do not reuse its
loops.  The rule, from KerComputeDisparity.cu:94-150: corners sorted by u
(A, B, C); column u belongs to the first part when A_u <= u < B_u and to the
second when B_u <= u < C_u (columns clipped to [0, W)); in its part, rows
[min(v_1, v_2), max(v_1, v_2)) are covered, v_1 on the line AC and v_2 on AB
or BC, each the (uint32_t) cast of the float line value; a part whose two
corners share a column is empty.  A covered pixel gets
(int32_t)(a*u + b*v + c) in float without contraction; a later triangle
overwrites an earlier one (the convention that fixes the reference's race).

Three independent checks:
  * hand-counted masks of triangles whose lines are exact in float;
  * a per-pixel membership predicate (one boolean per pixel and triangle,
    evaluated on the whole image at once) compared with the oracle on random
    lattice triangles, left and right (shifted) images;
  * geometry in float64: pixels more than one row inside a triangle's
    vertical extent are covered, pixels outside it are not.
An inclusive row end, a part boundary moved by one column or a rounded
(instead of truncated) span end fails at least one of them.
"""
import numpy as np
import pytest

STEP = 5


def _scene(W, H, corners):
    """Lattice and support points of triangles given as ((u, v, d), ...)
    triples; every corner gets its own support point."""
    can = np.full((H // STEP, W // STEP), -1, np.int32)
    sp, tri = [], []
    for t in corners:
        ids = []
        for (u, v, d) in t:
            assert u % STEP == 0 and v % STEP == 0
            can[v // STEP, u // STEP] = d
            ids.append(len(sp))
            sp.append((u, v))
        tri.append(ids)
    return can, np.array(sp, np.float32), np.array(tri, np.int32)


def _line(P, Q):
    """Slope and intercept of the line through P and Q, float32 as the
    reference computes them (slope 0 when both share a column)."""
    f = np.float32
    a = f(0)
    if int(P[0]) != int(Q[0]):
        a = f(f(P[1]) - f(Q[1])) / f(f(P[0]) - f(Q[0]))
    b = f(f(P[1]) - f(a * f(P[0])))
    return a, b


def _sat_u32_as_i32(x):
    """(int32_t)(uint32_t)x for float32 x, with the device's saturation."""
    x = np.where(np.isnan(x), 0, x).astype(np.float64)
    x = np.clip(np.floor(np.maximum(x, 0)), 0, 4294967295).astype(np.uint64)
    return x.astype(np.uint32).view(np.int32)


def _predicate_raster(can, sp, tri, right, W, H, planes):
    """Per-pixel form: for each triangle a boolean mask over the whole image
    from the membership rule above; later triangles overwrite."""
    out = np.full((H, W), -1.0, np.float32)
    U = np.arange(W, dtype=np.int64)[None, :].repeat(H, 0)
    V = np.arange(H, dtype=np.int64)[:, None].repeat(W, 1)
    Uf = U.astype(np.float32)
    Vf = V.astype(np.float32)
    for t, abc in zip(tri, planes):
        pts = []
        for i in t:
            u, v = int(sp[i, 0]), int(sp[i, 1])
            d = int(can[v // STEP, u // STEP])
            pts.append((u - d if right else u, v))
        # any order of corners with equal u gives the same spans (the two
        # lines of the non-empty part are the same pair either way)
        A, B, C = sorted(pts, key=lambda p: p[0])
        ac, ab, bc = _line(A, C), _line(A, B), _line(B, C)
        lac = _sat_u32_as_i32(ac[0] * Uf + ac[1])
        mask = np.zeros((H, W), bool)
        for lo, hi, other in ((A[0], B[0], ab), (B[0], C[0], bc)):
            if lo == hi:
                continue
            lo_ = _sat_u32_as_i32(other[0] * Uf + other[1])
            in_cols = (U >= lo) & (U < hi)
            in_rows = (V >= np.minimum(lac, lo_)) & (V < np.maximum(lac, lo_))
            mask |= in_cols & in_rows
        a, b, c = (np.float32(x) for x in abc)
        val = ((a * Uf) + (b * Vf)) + c
        val = np.trunc(val.astype(np.float64)).astype(np.float32)
        out = np.where(mask, val, out)
    return out


def _planes(oracle, can, sp, tri, right):
    return [oracle.plane(can, STEP, sp, t, right) for t in tri]


def test_hand_counted_masks(oracle):
    prm = oracle.params(disp_max=63)
    W, H = 40, 30
    # right triangle A=(0,0) B=(10,0) C=(10,10), constant d: column u in
    # [0, 10) covers rows [0, u) -> 1 + 2 + ... + 9 = 45 pixels
    can, sp, tri = _scene(W, H, [((0, 0, 7), (10, 0, 7), (10, 10, 7))])
    m = oracle.raster(can, sp, tri, 0, W, H, prm)
    cov = m >= 0
    assert cov.sum() == 45 and (m[cov] == 7).all()
    assert all(cov[v, u] == (v < u < 10) for u in range(W) for v in range(H))
    # apex triangle A=(0,5) B=(10,0) C=(20,5): first part rows [floor(5 - u/2), 5)
    # for u in [0, 10) -> 25 pixels, second part [floor(u/2 - 5), 5) for
    # u in [10, 20) -> 30 pixels; no pixel of row 5 (the half-open row end)
    can, sp, tri = _scene(W, H, [((0, 5, 3), (10, 0, 3), (20, 5, 3))])
    m = oracle.raster(can, sp, tri, 0, W, H, prm)
    cov = m >= 0
    assert cov[:, :10].sum() == 25 and cov[:, 10:20].sum() == 30 and cov.sum() == 55
    assert not cov[5:].any() and not cov[:, 20:].any()
    assert [int(cov[:, u].sum()) for u in range(20)] == [0, 1, 1, 2, 2, 3, 3, 4, 4, 5,
                                                          5, 5, 4, 4, 3, 3, 2, 2, 1, 1]
    # vertical left edge A=(5,0) B=(5,10) C=(15,5): first part empty (A_u == B_u);
    # second part u in [5, 15) between AC (v = u/2 - 2.5) and BC (v = 12.5 - u/2)
    can, sp, tri = _scene(W, H, [((5, 0, 2), (5, 10, 2), (15, 5, 2))])
    m = oracle.raster(can, sp, tri, 0, W, H, prm)
    cov = m >= 0
    exp = [int(np.floor(12.5 - u / 2) - np.floor(u / 2 - 2.5)) for u in range(5, 15)]
    assert [int(cov[:, u].sum()) for u in range(5, 15)] == exp
    assert cov.sum() == sum(exp) and not cov[:, :5].any() and not cov[:, 15:].any()


def test_overlap_highest_index_wins(oracle):
    prm = oracle.params(disp_max=63)
    W, H = 40, 30
    t1 = ((0, 0, 7), (20, 0, 7), (20, 20, 7))
    t2 = ((5, 0, 11), (25, 0, 11), (25, 20, 11))
    # t1 covers rows [0, u) of columns [0, 20), t2 rows [0, u - 5) of [5, 25)
    U, V = np.meshgrid(np.arange(W), np.arange(H))
    only1 = (V < U) & (U < 20) & ~((V < U - 5) & (U >= 5))
    only2 = (V < U - 5) & (U >= 20) & (U < 25)
    both = (V < U - 5) & (U >= 5) & (U < 20)
    for order, top in (((t1, t2), 11), ((t2, t1), 7)):
        can, sp, tri = _scene(W, H, list(order))
        m = oracle.raster(can, sp, tri, 0, W, H, prm)
        assert both.sum() == 105 and (m[both] == top).all()
        assert (m[only1] == 7).all() and (m[only2] == 11).all()
        assert ((m >= 0) == (only1 | only2 | both)).all()


def _random_triangles(rng, W, H, n, dmax):
    out = []
    for _ in range(n):
        cu = rng.integers(0, W // STEP, 3) * STEP
        cv = rng.integers(0, H // STEP, 3) * STEP
        if len(set(zip(cu, cv))) < 3:
            continue
        d = rng.integers(0, dmax, 3)
        out.append(tuple((int(cu[k]), int(cv[k]), int(d[k])) for k in range(3)))
    return out


def _disjoint(tris):
    """Keep triangles whose corners are not shared with earlier ones (each
    lattice point carries one disparity)."""
    seen, keep = set(), []
    for t in tris:
        pts = {(u, v) for u, v, _ in t}
        if pts & seen:
            continue
        seen |= pts
        keep.append(t)
    return keep


@pytest.mark.parametrize("right", [0, 1])
def test_oracle_equals_per_pixel_predicate(oracle, right):
    prm = oracle.params(disp_max=63)
    rng = np.random.default_rng(11 + right)
    W, H = 90, 70
    for rep in range(8):
        tris = _disjoint(_random_triangles(rng, W, H, 60, 40))
        can, sp, tri = _scene(W, H, tris)
        planes = _planes(oracle, can, sp, tri, right)
        got = oracle.raster(can, sp, tri, right, W, H, prm)
        exp = _predicate_raster(can, sp, tri, right, W, H, planes)
        assert np.array_equal(got, exp), (rep, int((got != exp).sum()))
        assert (got >= 0).sum() > 200


def test_predicate_equals_oracle_on_delaunay_meshes(oracle, synth):
    """The same on the meshes of a real pipeline run (shared corners,
    left and right images)."""
    left, right_img, _ = synth.make_pair(160, 120, 31, regions=4, seed=5)
    prm = oracle.params(disp_max=31)
    out = oracle.process(left, right_img, prm)
    can = out["can_final"]
    for right, tri, mk in ((0, out["tri0"], "mean0"), (1, out["tri1"], "mean1")):
        H, W = left.shape
        planes = _planes(oracle, can, out["sp0"], tri, right)
        exp = _predicate_raster(can, out["sp0"], tri, right, W, H, planes)
        assert np.array_equal(out[mk], exp)
        assert np.array_equal(oracle.raster(can, out["sp0"], tri, right, W, H, prm), exp)


@pytest.mark.parametrize("right", [0, 1])
def test_geometric_interior_and_exterior(oracle, right):
    """float64 geometry, independent of the span formula: at column u a
    triangle covers the closed interval [lo(u), hi(u)] of v.  Rows v with
    lo + 1 <= v <= hi - 1 for A_u < u < C_u must be covered (floor(lo) <= v
    < floor(hi)), rows v < lo - 1, v > hi and columns outside [A_u, C_u) must
    not.  The margin of one row is the float rounding of the span ends: a line
    through an integer row can evaluate just below it (12.999999 for 13),
    which the truncation then moves down a whole row."""
    prm = oracle.params(disp_max=63)
    rng = np.random.default_rng(23 + right)
    W, H = 90, 70
    n_in = n_out = 0
    for rep in range(20):
        tris = _random_triangles(rng, W, H, 1, 40)
        if not tris:
            continue
        can, sp, tri = _scene(W, H, tris)
        got = oracle.raster(can, sp, tri, right, W, H, prm) >= 0
        pts = [(u - d if right else u, v) for u, v, d in tris[0]]
        us = [p[0] for p in pts]
        for u in range(W):
            ys = []
            for (p, q) in ((pts[0], pts[1]), (pts[1], pts[2]), (pts[0], pts[2])):
                if p[0] == q[0]:
                    if p[0] == u:
                        ys += [p[1], q[1]]
                elif min(p[0], q[0]) <= u <= max(p[0], q[0]):
                    ys.append(p[1] + (q[1] - p[1]) * (u - p[0]) / (q[0] - p[0]))
            for v in range(H):
                if not (min(us) <= u < max(us)) or not ys:
                    assert not got[v, u]
                    continue
                lo, hi = min(ys), max(ys)
                if min(us) < u and lo + 1 + 1e-6 <= v <= hi - 1 - 1e-6:
                    assert got[v, u], (tris, u, v)
                    n_in += 1
                if v < lo - 1 - 1e-6 or v > hi + 1e-6:
                    assert not got[v, u], (tris, u, v)
                    n_out += 1
    assert n_in > 100 and n_out > 1000
