"""Dense matching under the tilings the autotuner chooses from.

Every tiling (cells per wave x steps per chunk) must give the oracle's dense
disparities bit for bit; the autotuner records one tiling per shape; a batch
whose pairs use different tilings (two dense launches in one chunk) equals
the plain calls.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TILINGS = [(1, 1), (1, 2), (1, 3), (1, 8), (2, 1), (2, 2), (2, 4)]


def _segs(W, cpw, gs=20):
    gw = -(-W // gs)
    return -(-gw // (8 * cpw))


@pytest.mark.parametrize("name", ["450x375_d64", "1242x375_d128"])
def test_every_tiling_bit_exact(pkg, oracle, pairs, name):
    left, right, _, dmax = pairs(name)
    H, W = left.shape
    ref = oracle.process(left, right, oracle.params(disp_max=dmax))
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    el.set_debug(True)
    for cpw, spc in TILINGS:
        print("tiling", name, cpw, spc, flush=True)
        el.set_dense_tiles(cpw, spc)
        assert el.dense_tiles(W, H, dmax)[:2] == (cpw, min(spc, _segs(W, cpw)))
        D1, D2 = el.process(left, right)
        assert el.status == 0
        for k in ("d1_raw", "d2_raw"):
            got = el.fetch(k)
            if not np.array_equal(got, ref[k]):
                bad = np.argwhere(got != ref[k])
                raise AssertionError("%s tiling %s: %s differs at %d px, first %s (gpu %s, oracle %s)" % (
                    name, (cpw, spc), k, len(bad), bad[0].tolist(), got[tuple(bad[0])], ref[k][tuple(bad[0])]))
        assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])


def test_large_pair_tilings_bit_exact(pkg, oracle, pairs):
    """2816x1880, disparities 0..255: the rings wrap many times per chunk."""
    left, right, _, dmax = pairs("2816x1880_d256")
    ref = oracle.process(left, right, oracle.params(disp_max=dmax), dump=False)
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    for cpw, spc in [(1, 1), (1, 5), (2, 3), (2, 9)]:
        print("tiling 2816x1880", cpw, spc, flush=True)
        el.set_dense_tiles(cpw, spc)
        D1, D2 = el.process(left, right)
        assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"]), (cpw, spc)


def test_autotune_records_a_tiling_per_shape(pkg, pairs):
    left, right, _, dmax = pairs("1242x375_d128")
    H, W = left.shape
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    assert el.dense_tiles(W, H, dmax)[2] == -1.0
    a1, a2 = el.process(left, right)
    cpw, spc, us = el.dense_tiles(W, H, dmax)
    print("tuned", (W, H, dmax), cpw, spc, us, flush=True)
    assert cpw in (1, 2) and 1 <= spc <= _segs(W, cpw) and us > 0
    b1, b2 = el.process(left, right)  # runs with the tuned tiling
    assert np.array_equal(a1, b1) and np.array_equal(a2, b2)
    off = pkg.Elas(pkg.default_params(disp_max=dmax))
    off.set_autotune(False)
    c1, c2 = off.process(left, right)
    assert off.dense_tiles(W, H, dmax)[2] == -1.0
    assert np.array_equal(a1, c1) and np.array_equal(a2, c2)


def test_batch_mixing_tilings_equals_oracle(pkg, oracle, synth):
    """disp_max 383 does not fit two cells per wave in 64 KiB of LDS at grid
    size 20 (one cell per wave); 31 and 63 do (two): one chunk holds both
    dense launches."""
    shapes = [(600, 100, 383), (200, 90, 63), (520, 110, 383), (160, 120, 31)]
    pairs_ = [synth.make_pair(W, H, dm, regions=3, seed=500 + i)[:2] + (dm,) for i, (W, H, dm) in enumerate(shapes)]
    el = pkg.Elas(pkg.default_params(disp_max=383))
    el.set_autotune(False)
    assert el.dense_tiles(600, 100, 383)[0] == 1 and el.dense_tiles(200, 90, 63)[0] == 2
    batch = el.process_batch(pairs_)
    for i, ((l, r, dm), (b1, b2, st)) in enumerate(zip(pairs_, batch)):
        print("mixed batch pair", i, l.shape, dm, flush=True)
        ref = oracle.process(l, r, oracle.params(disp_max=dm), dump=False)
        assert st == ref["status"]
        assert np.array_equal(b1, ref["D1"]) and np.array_equal(b2, ref["D2"])


@pytest.mark.parametrize("name", ["1242x375_d128", "2816x1880_d256"])
def test_every_descriptor_strip_height_bit_exact(pkg, oracle, pairs, name):
    """Strips of 6, 12 and 18 rows give the oracle's descriptors; a shape's
    tuned height is recorded with its time."""
    left, right, _, dmax = pairs(name)
    H, W = left.shape
    d0, d1 = oracle.descriptor(left), oracle.descriptor(right)
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    el.set_debug(True)
    for rows in (6, 12, 18):
        print("descriptor rows", name, rows, flush=True)
        el.set_desc_rows(rows)
        assert el.desc_rows(W, H, dmax)[0] == rows
        el.process(left, right)
        assert np.array_equal(el.fetch("desc0"), d0) and np.array_equal(el.fetch("desc1"), d1), rows
    with pytest.raises(pkg.ElasError):
        el.set_desc_rows(7)
    el.set_desc_rows(0)
    tuned = pkg.Elas(pkg.default_params(disp_max=dmax))
    tuned.process(left, right)
    rows, us = tuned.desc_rows(W, H, dmax)
    print("tuned descriptor rows", name, rows, us, flush=True)
    assert rows in (6, 12, 18) and us > 0
