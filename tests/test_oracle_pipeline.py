"""Whole-pipeline properties of the oracle (src/Elas.cu:19-120) and of the
synthetic generator; the golden fixture pins the oracle against regressions."""
import hashlib
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_generator_is_pinned(synth):
    """SHA-256 of the seeded 450x375 pair (numpy PCG64 + float64 arithmetic)."""
    l, r, _, _ = synth.config_pair("450x375_d64")
    pins = open(os.path.join(GOLDEN, "synth_sha256.txt")).read().split()
    assert hashlib.sha256(l.tobytes()).hexdigest() == pins[0]
    assert hashlib.sha256(r.tobytes()).hexdigest() == pins[1]


def test_pipeline_recovers_ground_truth(oracle, pairs):
    left, right, gt, dmax = pairs("450x375_d64")
    out = oracle.process(left, right, oracle.params(disp_max=dmax))
    D1, D2 = out["D1"], out["D2"]
    v = D1 >= 0
    assert out["status"] == 0 and v.mean() > 0.75
    assert (np.abs(D1[v] - gt[v]) <= 1).mean() > 0.95
    for D in (D1, D2):
        valid = D[D >= 0]
        assert ((D == -1) | (D >= 0)).all() and (valid == np.round(valid)).all() and valid.max() <= dmax


def test_final_maps_are_lr_consistent(oracle, pairs):
    left, right, _, dmax = pairs("450x375_d64")
    out = oracle.process(left, right, oracle.params(disp_max=dmax))
    D1, D2 = out["D1"], out["D2"]
    r1, r2 = out["d1_raw"], out["d2_raw"]
    H, W = D1.shape
    for v in range(0, H, 7):
        for u in range(W):
            d = D1[v, u]
            if d >= 0:
                assert abs(r2[v, int(u - d)] - d) <= 2
            d = D2[v, u]
            if d >= 0:
                assert abs(r1[v, int(u + d)] - d) <= 2


def _incircle_violations(pts, tri):
    """Points strictly inside a triangle's circumcircle (exact integers)."""
    P = pts.astype(np.int64)
    bad = 0
    for t in tri:
        a, b, c = P[t[0]], P[t[1]], P[t[2]]
        orient = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
        d = P - 0
        adx, ady = a[0] - d[:, 0], a[1] - d[:, 1]
        bdx, bdy = b[0] - d[:, 0], b[1] - d[:, 1]
        cdx, cdy = c[0] - d[:, 0], c[1] - d[:, 1]
        det = ((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy)
               - (bdx * bdx + bdy * bdy) * (adx * cdy - cdx * ady)
               + (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady))
        bad += int(((det * np.sign(orient)) > 0).sum())
    return bad


def test_triangle_meshes_are_delaunay(oracle, pairs):
    left, right, _, dmax = pairs("450x375_d64")
    out = oracle.process(left, right, oracle.params(disp_max=dmax))
    for sp, tri in ((out["sp0"], out["tri0"]), (out["sp1"], out["tri1"])):
        assert len(tri) > len(sp)
        assert _incircle_violations(sp, tri) == 0


def test_oracle_is_deterministic(oracle, pairs):
    left, right, _, dmax = pairs("1242x375_d128")
    a = oracle.process(left, right, oracle.params(disp_max=dmax), dump=False)
    b = oracle.process(left, right, oracle.params(disp_max=dmax), dump=False)
    assert np.array_equal(a["D1"], b["D1"]) and np.array_equal(a["D2"], b["D2"])


def test_golden_fixture(oracle):
    """tests/golden/small_pair.npz was written by tests/golden/make_golden.py."""
    g = np.load(os.path.join(GOLDEN, "small_pair.npz"))
    out = oracle.process(g["left"], g["right"], oracle.params(disp_max=int(g["disp_max"])))
    for k in ("D1", "D2", "sp0", "tri0", "tri1", "can_final"):
        assert np.array_equal(out[k], g[k]), k


def test_too_few_points_is_an_error_not_an_exit(oracle):
    img = np.full((50, 60), 3, np.uint8)
    out = oracle.process(img, img, oracle.params(disp_max=15))
    assert out["status"] == 4 and (out["D1"] == -1).all()
    with pytest.raises(ValueError):
        oracle.triangulate(np.zeros((2, 2), np.float32))
