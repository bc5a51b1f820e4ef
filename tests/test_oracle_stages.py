"""Pins of the oracle's later stages against independent formulations:
filters (KerSupportMatches.cu:196-356), compaction (:361-423), grid
(KerGrid.cu), plane (KerComputeDisparity.cu:10-124,322-375), dense argmin
(:156-316), LR check (KerLeftRightConsistencyCheck.cu)."""
from fractions import Fraction

import numpy as np
import pytest


def _random_can(rng, ch, cw, p_valid=0.5, dmax=40):
    can = rng.integers(0, dmax, (ch, cw)).astype(np.int32)
    can[rng.random((ch, cw)) > p_valid] = -1
    return can


def test_incon_filter_by_window_counts(oracle):
    """Jacobi: count of valid neighbours within +-R (both axes) with
    |d - d2| <= thr, computed by shifting whole arrays."""
    rng = np.random.default_rng(0)
    prm = oracle.params()
    for ch, cw in [(20, 30), (31, 17)]:  # the second is portrait: rows >= cw are not filtered
        can = _random_can(rng, ch, cw, 0.6, 20)
        R = prm.incon_window_size
        pad = np.full((ch + 2 * R, cw + 2 * R), -1, np.int64)
        pad[R:R + ch, R:R + cw] = can
        support = np.zeros((ch, cw), np.int64)
        for dy in range(-R, R + 1):
            for dx in range(-R, R + 1):
                nb = pad[R + dy:R + dy + ch, R + dx:R + dx + cw]
                support += (nb >= 0) & (np.abs(nb - can) <= prm.incon_threshold)
        exp = can.copy()
        rows = np.arange(ch)[:, None] < cw
        exp[(can >= 0) & (support < prm.incon_min_support) & rows] = -1
        assert np.array_equal(oracle.incon_filter(can, prm), exp)


def _near(a, d):
    return a >= 0 and abs(int(a) - int(d)) <= 1


def test_redundancy_filter_properties(oracle):
    """Horizontal then vertical sequential scans.  Checked per line: a cell is
    removed iff, at its turn, a valid cell within 5 on the already-scanned side
    (final values) and on the other side (values before this scan) is within 1."""
    rng = np.random.default_rng(1)
    for ch, cw in [(15, 40), (40, 15), (1, 30), (30, 1)]:
        can = _random_can(rng, ch, cw, 0.7, 6)
        out = oracle.redun_filter(can)
        h = can.copy()
        for v in range(ch):        # reconstruct the horizontal result from the rule
            for u in range(cw):
                d = can[v, u]
                if d < 0:
                    continue
                left = any(_near(h[v, j], d) for j in range(max(0, u - 5), u))
                right = any(_near(can[v, j], d) for j in range(u + 1, min(cw, u + 6)))
                if left and right:
                    h[v, u] = -1
        exp = h.copy()
        for u in range(cw):
            for v in range(ch):
                d = h[v, u]
                if d < 0:
                    continue
                up = any(_near(exp[j, u], d) for j in range(max(0, v - 5), v))
                down = any(_near(h[j, u], d) for j in range(v + 1, min(ch, v + 6)))
                if up and down:
                    exp[v, u] = -1
        assert np.array_equal(out, exp)
        assert ((out >= 0) <= (can >= 0)).all()


def test_compaction_is_row_major(oracle):
    rng = np.random.default_rng(2)
    can = _random_can(rng, 13, 21, 0.3)
    sp0, sp1 = oracle.compact(can, 5)
    vs, us = np.nonzero(can >= 0)
    assert np.array_equal(sp0, np.stack([us * 5, vs * 5], 1).astype(np.float32))
    assert np.array_equal(sp1, np.stack([us * 5 - can[vs, us], vs * 5], 1).astype(np.float32))


def test_grid_as_sets(oracle):
    """Cell sets: d-1..d+1 (clamped) of every support point, then the 3x3 union
    for interior cells; borders empty; lists ascending with a count."""
    rng = np.random.default_rng(3)
    W, H, dmax = 260, 140, 63
    prm = oracle.params(disp_max=dmax)
    can = _random_can(rng, H // 5, W // 5, 0.2, 50)
    can[can >= 0] = np.minimum(can[can >= 0], np.arange(W // 5)[None, :].repeat(H // 5, 0)[can >= 0] * 5)
    g0, g1 = oracle.grid(can, W, H, prm)
    gw, gh = -(-W // 20), -(-H // 20)
    sets = [[[set() for _ in range(gw)] for _ in range(gh)] for _ in range(2)]
    for vc, uc in zip(*np.nonzero(can >= 0)):
        d = int(can[vc, uc])
        for k, u in enumerate((uc * 5, uc * 5 - d)):
            gx, gy = u // 20, (vc * 5) // 20
            if 0 <= gx < gw and 0 <= gy < gh:
                sets[k][gy][gx] |= set(range(max(d - 1, 0), min(d + 1, dmax) + 1))
    for k, g in enumerate((g0, g1)):
        for gy in range(gh):
            for gx in range(gw):
                cell = g[gy, gx]
                if 0 < gx < gw - 1 and 0 < gy < gh - 1:
                    s = set().union(*[sets[k][gy + a][gx + b] for a in (-1, 0, 1) for b in (-1, 0, 1)])
                else:
                    s = set()
                n = cell[0]
                assert n == len(s) and list(cell[1:1 + n]) == sorted(s) and not cell[1 + n:].any()


def test_plane_equals_exact_rational_plane(oracle, pairs):
    """Gauss-Jordan in double, cast to float, agrees with the exact plane
    through three lattice corners (Cramer's rule over the rationals) to float
    precision; most coefficients are bit-identical (double rounding can leave
    e.g. 5.6e-17 where the exact value is 0)."""
    left, right, _, dmax = pairs("450x375_d64")
    out = oracle.process(left, right, oracle.params(disp_max=dmax))
    can, sp0 = out["can_final"], out["sp0"]
    n_same = n_total = 0
    for right_img, tri in ((0, out["tri0"]), (1, out["tri1"])):
        for t in tri[:: max(1, len(tri) // 300)]:
            pts = []
            for i in t:
                u, v = int(sp0[i, 0]), int(sp0[i, 1])
                d = int(can[v // 5, u // 5])
                pts.append((u - d if right_img else u, v, d))
            (u0, v0, d0), (u1, v1, d1), (u2, v2, d2) = pts
            det = Fraction(u0 * (v1 - v2) - v0 * (u1 - u2) + (u1 * v2 - u2 * v1))
            a = Fraction(d0 * (v1 - v2) - v0 * (d1 - d2) + (d1 * v2 - d2 * v1)) / det
            b = Fraction(u0 * (d1 - d2) - d0 * (u1 - u2) + (u1 * d2 - u2 * d1)) / det
            c = Fraction(u0 * (v1 * d2 - v2 * d1) - v0 * (u1 * d2 - u2 * d1) + d0 * (u1 * v2 - u2 * v1)) / det
            got = oracle.plane(can, 5, sp0, t, right_img)
            exp = np.array([float(a), float(b), float(c)], np.float32)
            assert np.allclose(got, exp, rtol=2e-7, atol=1e-12), (t, got, exp)
            n_total += 3
            n_same += int((got == exp).sum())
    assert n_same > 0.5 * n_total


def test_constant_disparity_scene_gives_constant_prior(oracle):
    """All support points at d = 9: every covered pixel's prior is exactly 9."""
    W, H = 120, 90
    can = np.full((H // 5, W // 5), -1, np.int32)
    can[2:-2:3, 3:-2:4] = 9
    sp0, _ = oracle.compact(can, 5)
    prm = oracle.params(disp_max=31)
    tri = oracle.triangulate(sp0)
    mean = oracle.raster(can, sp0, tri, 0, W, H, prm)
    cov = mean >= 0
    assert cov.mean() > 0.4 and (mean[cov] == 9).all()


def _dense_key_form(grid, desc0, desc1, mean, right, prm):
    """Dense argmin as the minimum of the key (cost, phase, d): phase 0 = grid
    candidates outside the window, 1 = window with prior.  Independent of the
    reference's loop structure."""
    H, W = mean.shape
    out = np.full((H, W), -1, np.float32)
    I1, I2 = (desc1, desc0) if right else (desc0, desc1)
    I1 = I1.astype(np.int32)
    I2 = I2.astype(np.int32)
    pr = oracle_pr = 2
    P = {0: -14, 1: -9, 2: -2}
    for v in range(4, H - 3):
        for u in range(4, W - 3):
            dp = int(mean[v, u])
            if dp <= 0:
                continue
            lo, hi = max(dp - pr, 0), min(dp + pr, prm.disp_max)
            cell = grid[v // 20, u // 20]
            cands = [(0, int(d)) for d in cell[1:1 + cell[0]] if not lo <= d <= hi]
            cands += [(1, d) for d in range(lo, hi + 1)]
            best = None
            for ph, d in cands:
                uw = u + d if right else u - d
                if not 3 <= uw <= W - 4:
                    continue
                cost = int(np.abs(I2[v, uw] - I1[v, u]).sum()) + (P[abs(d - dp)] if ph else 0)
                key = (cost, ph, d)
                best = key if best is None or key < best else best
            if best is not None:
                out[v, u] = best[2]
    return out


def test_dense_matching_key_form(oracle, synth):
    W, H, dmax = 90, 50, 31
    l, r, _ = synth.make_pair(W, H, dmax, regions=3, seed=21)
    prm = oracle.params(disp_max=dmax)
    out = oracle.process(l, r, prm)
    for right, gk, mk, dk in ((0, "grid0", "mean0", "d1_raw"), (1, "grid1", "mean1", "d2_raw")):
        exp = _dense_key_form(out[gk], out["desc0"], out["desc1"], out[mk], right, prm)
        got, _ = oracle.dense(out[gk], out["desc0"], out["desc1"], out[mk], right, prm)
        assert np.array_equal(got, exp)
        assert np.array_equal(got, out[dk])
        assert (got >= 0).sum() > 500


def test_lr_check_vectorised(oracle):
    rng = np.random.default_rng(4)
    H, W = 30, 60
    D1 = rng.integers(-1, 12, (H, W)).astype(np.float32)
    D2 = rng.integers(-1, 12, (H, W)).astype(np.float32)
    o1, o2 = oracle.lr_check(D1, D2, oracle.params())
    u = np.arange(W)[None, :].repeat(H, 0)
    v = np.arange(H)[:, None].repeat(W, 1)
    for D, Do, o, sgn in ((D1, D2, o1, -1), (D2, D1, o2, 1)):
        uw = u + sgn * D
        ok = (D >= 0) & (uw >= 0) & (uw < W)
        partner = np.where(ok, Do[v, np.clip(uw, 0, W - 1).astype(int)], 0)
        keep = ok & (np.abs(partner - D) <= 2)
        assert np.array_equal(o, np.where(keep, D, -1))
