"""Support matching (KerSupportMatches.cu:13-192): the oracle simulates the
reference's 32 threads and thread 0's serial reduction literally; these tests
check it against a closed-form statement computed from the full cost table,
and against shift invariants."""
import numpy as np
import pytest


def _closed_form_pass(desc0, desc1, prm, right, can_in):
    """Decision rule in closed form.  With E(d) the 64-byte cost, m = min E,
    (r*, d*) the lexicographic minimum of ((d - dmin) % 32, d) over the
    minimisers and m2 the second smallest E counted with multiplicity: thread 0
    holds residue 0 and its best is counted twice, so min2 = m when r* == 0,
    else m2; accept when (float)m < thr * (float)min2, winner d*."""
    H, W = desc0.shape[:2]
    step = prm.candidate_stepsize
    cw, ch = W // step, H // step
    can = can_in.copy()
    offs = [(-2, -2), (2, -2), (-2, 2), (2, 2)]
    thr = np.float32(prm.support_threshold)
    D0 = desc0.astype(np.int32)
    D1 = desc1.astype(np.int32)
    for vc in range(ch):
        for uc in range(cw):
            u, v = uc * step, vc * step
            if not (5 <= u <= W - 6 and 5 <= v <= H - 6):
                continue
            dl = 0
            if right:
                dl = can[vc, uc]
                if dl < 0:
                    continue
                u -= dl
            I1, I2 = (D1, D0) if right else (D0, D1)
            if np.abs(I1[v, u] - 128).sum() < prm.support_texture:
                can[vc, uc] = -1
                continue
            dmin = max(prm.disp_min, 0)
            dmax = min(prm.disp_max, (W - u - 5) if right else (u - 5))
            if dmax - dmin < 10:
                can[vc, uc] = -1
                continue
            ds = np.arange(dmin, dmax + 1)
            uw = u + ds if right else u - ds
            E = np.zeros(len(ds), np.int64)
            for x, y in offs:
                E += np.abs(I2[v + y, uw + x] - I1[v + y, u + x][None, :]).sum(axis=1)
            m = E.min()
            mins = ds[E == m]
            key = sorted(((d - dmin) % 32, d) for d in mins)[0]
            m2 = np.sort(E)[1]
            min2 = m if key[0] == 0 else m2
            if np.float32(m) < thr * np.float32(min2):
                if not right:
                    can[vc, uc] = key[1]
                elif abs(dl - key[1]) > prm.lr_threshold:
                    can[vc, uc] = -1
            else:
                can[vc, uc] = -1
    return can


@pytest.mark.parametrize("seed,thr", [(0, 0.85), (1, 0.85), (2, 0.95), (3, 1.5), (4, 2.5)])
def test_support_pass_equals_closed_form(oracle, synth, seed, thr):
    """thr > 1 accepts ties and exercises which minimiser wins."""
    W, H, dmax = 96, 40, 47
    l, r, _ = synth.make_pair(W, H, dmax, regions=3, seed=seed)
    if seed >= 3:  # quantise to create many cost ties
        l = (l // 64 * 64).astype(np.uint8)
        r = (r // 64 * 64).astype(np.uint8)
    prm = oracle.params(disp_max=dmax, support_threshold=thr)
    d0, d1 = oracle.descriptor(l), oracle.descriptor(r)
    can = np.full((H // 5, W // 5), -1, np.int32)
    oracle.support_pass(d0, d1, prm, 0, can)
    exp_l = _closed_form_pass(d0, d1, prm, 0, np.full_like(can, -1))
    assert np.array_equal(can, exp_l)
    can_r = can.copy()
    oracle.support_pass(d0, d1, prm, 1, can_r)
    assert np.array_equal(can_r, _closed_form_pass(d0, d1, prm, 1, can))
    assert (can >= 0).sum() > 0


def _shifted_pair(k, W=160, H=60, seed=5):
    rng = np.random.default_rng(seed)
    big = rng.integers(0, 256, (H, W + k), dtype=np.uint8)
    left = big[:, :W].copy()
    right = big[:, k:k + W].copy()       # right(x) = left(x + k): disparity k
    return left, right


@pytest.mark.parametrize("k", [3, 7, 19, 31, 33, 45])
def test_integer_shift_gives_its_disparity(oracle, k):
    left, right = _shifted_pair(k)
    prm = oracle.params(disp_max=63)
    d0, d1 = oracle.descriptor(left), oracle.descriptor(right)
    can = np.full((60 // 5, 160 // 5), -1, np.int32)
    oracle.support_pass(d0, d1, prm, 0, can)
    reach = (np.arange(can.shape[1]) * 5 - 5 >= k)[None, :].repeat(can.shape[0], 0)
    vals = can[(can >= 0) & reach]            # candidates whose range [0, u-5] holds k
    assert len(vals) > 20 and (vals == k).all()


@pytest.mark.parametrize("k", [32])
def test_shift_by_multiple_of_32_is_rejected(oracle, k):
    """The unique best d = 32 sits in thread 0's residue class: thread 0's best
    is counted twice, min2 = min1 and the ratio test fails (Q1)."""
    left, right = _shifted_pair(k)
    prm = oracle.params(disp_max=63)
    d0, d1 = oracle.descriptor(left), oracle.descriptor(right)
    can = np.full((60 // 5, 160 // 5), -1, np.int32)
    oracle.support_pass(d0, d1, prm, 0, can)
    reach = (np.arange(can.shape[1]) * 5 - 5 >= k)[None, :].repeat(can.shape[0], 0)
    assert (can[reach] == -1).all()
    # the same scene shifted by k + 1 is matched: the rejection is the quirk, not the scene
    left, right = _shifted_pair(k + 1)
    can2 = np.full_like(can, -1)
    oracle.support_pass(oracle.descriptor(left), oracle.descriptor(right), prm, 0, can2)
    assert ((can2 == k + 1) & reach).sum() > 20


def test_textureless_candidates_rejected(oracle):
    img = np.full((40, 80), 50, np.uint8)
    d = oracle.descriptor(img)
    can = np.full((8, 16), 5, np.int32)
    oracle.support_pass(d, d, oracle.params(disp_max=31), 0, can)
    # in-bounds cells become -1; out-of-bounds cells keep their value
    inb = np.zeros_like(can, bool)
    for vc in range(8):
        for uc in range(16):
            inb[vc, uc] = 5 <= uc * 5 <= 80 - 6 and 5 <= vc * 5 <= 40 - 6
    assert (can[inb] == -1).all() and (can[~inb] == 5).all()
