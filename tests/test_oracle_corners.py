"""add_corners (src/Elas.h:23, declared by the reference without code): the
oracle's definition pinned by a brute-force nearest-neighbour search, and
its effect on the pipeline.  Subsampling is refused (no semantics)."""
import numpy as np
import pytest


def _corners_brute(sp0, sp1, W, H):
    d = (sp0[:, 0] - sp1[:, 0]).astype(np.int64)
    out0, out1 = [], []
    cd = []
    for cu, cv in ((0, 0), (0, H - 1), (W - 1, 0), (W - 1, H - 1)):
        dist = (sp0[:, 0].astype(np.int64) - cu) ** 2 + (sp0[:, 1].astype(np.int64) - cv) ** 2
        k = int(np.argmin(dist))               # first minimum in compaction order
        cd.append(int(d[k]))
        out0.append((cu, cv))
        out1.append((cu - d[k], cv))
    for c, v in ((2, 0), (3, H - 1)):         # right corners again at u = W-1+d
        out0.append((W - 1 + cd[c], v))
        out1.append((W - 1, v))
    return np.array(out0, np.float32), np.array(out1, np.float32)


def test_corner_points_are_nearest_neighbours(oracle, synth):
    for W, H, dm, seed in ((160, 120, 31, 3), (97, 131, 15, 4), (300, 90, 63, 5)):
        l, r, _ = synth.make_pair(W, H, dm, regions=4, seed=seed)
        out = oracle.process(l, r, oracle.params(disp_max=dm))
        sp0, sp1 = out["sp0"], out["sp1"]
        a, b = oracle.add_corners(sp0, sp1, W, H)
        e0, e1 = _corners_brute(sp0, sp1, W, H)
        assert np.array_equal(a[:len(sp0)], sp0) and np.array_equal(b[:len(sp1)], sp1)
        assert np.array_equal(a[len(sp0):], e0) and np.array_equal(b[len(sp1):], e1)


def test_ties_take_the_first_support_point(oracle):
    # (0,0)'s nearest: (5,5) and... two points at equal distance 50 from (0,0)
    sp0 = np.array([[10, 0], [5, 5], [0, 10], [50, 40]], np.float32)
    sp1 = sp0 - np.array([[3, 0], [7, 0], [9, 0], [2, 0]], np.float32)
    a, b = oracle.add_corners(sp0, sp1, 60, 50)
    # squared distances from (0,0): 100, 50, 100, 4100 -> index 1 (d = 7)
    assert tuple(a[4]) == (0, 0) and tuple(b[4]) == (-7, 0)
    # from (0, 49): 2501, 1961, 1521, 2581 -> index 2 (d = 9)
    assert tuple(b[5]) == (-9, 49)
    # from (59, 0) and (59, 49): index 3 both (d = 2); right corners again at 59 + 2
    assert tuple(b[6]) == (57, 0) and tuple(b[7]) == (57, 49)
    assert tuple(a[8]) == (61, 0) and tuple(b[8]) == (59, 0)
    assert tuple(a[9]) == (61, 49) and tuple(b[9]) == (59, 49)
    # with equal distances the first point wins: (0,0) -> [10,0] and [0,10] both 100
    sp0b = np.array([[10, 0], [0, 10]], np.float32)
    a, b = oracle.add_corners(sp0b, sp0b - [[4, 0], [6, 0]], 30, 30)
    assert tuple(b[2]) == (-4, 0)


def test_pipeline_with_corners(oracle, synth):
    W, H, dm = 200, 150, 31
    l, r, _ = synth.make_pair(W, H, dm, regions=4, seed=9)
    off = oracle.process(l, r, oracle.params(disp_max=dm))
    on = oracle.process(l, r, oracle.params(disp_max=dm, add_corners=True))
    assert on["nsp"] == off["nsp"] + 6
    assert np.array_equal(on["sp0"][:off["nsp"]], off["sp0"])
    # the corners extend both meshes to the image corners: every corner of the
    # left prior is covered or on the half-open span ends, and more of both
    # priors is covered than without them
    for k in ("mean0", "mean1"):
        assert (on[k] >= 0).sum() > (off[k] >= 0).sum()
    assert (on["mean0"][1:-1, 1:-1] >= 0).mean() > 0.98
    assert on["status"] == 0 and (on["D1"] >= 0).mean() > (off["D1"] >= 0).mean() - 0.01


def test_subsampling_is_refused(pkg):
    with pytest.raises(pkg.ElasError) as e:
        pkg.Elas(pkg.default_params(subsampling=True), device=0)
    assert e.value.code == 6        # refused before any device is touched
