"""add_corners on the GPU: the corner support points, meshes, priors and
maps equal the oracle's, alone and in a batch; subsampling is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STAGE_KEYS = ["sp0", "sp1", "tri0", "tri1", "mean0", "mean1", "d1_raw", "d2_raw"]


@pytest.mark.parametrize("name", ["450x375_d64", "1242x375_d128"])
def test_corners_every_stage_bit_exact(pkg, oracle, pairs, name):
    left, right, _, dmax = pairs(name)
    print("case", name, "add_corners", flush=True)
    el = pkg.Elas(pkg.default_params(disp_max=dmax, add_corners=True))
    el.set_debug(True)
    D1, D2 = el.process(left, right)
    ref = oracle.process(left, right, oracle.params(disp_max=dmax, add_corners=True))
    assert el.status == ref["status"] == 0
    assert el.shape_info()["num_support"] == ref["nsp"]
    for k in STAGE_KEYS:
        got, exp = el.fetch(k), ref[k]
        assert got.shape == exp.shape and np.array_equal(got, exp), k
    assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])
    el.close()


def test_corners_in_a_batch(pkg, oracle, synth):
    shapes = [(160, 120, 31), (200, 90, 63), (97, 131, 15), (240, 180, 63)]
    pairs_ = [synth.make_pair(W, H, dm, regions=4, seed=800 + i)[:2] + (dm,) for i, (W, H, dm) in enumerate(shapes)]
    el = pkg.Elas(pkg.default_params(disp_max=63, add_corners=True))
    el.set_pipeline(2, 2)
    out = el.process_batch(pairs_)
    for (l, r, dm), (d1, d2, st) in zip(pairs_, out):
        ref = oracle.process(l, r, oracle.params(disp_max=dm, add_corners=True), dump=False)
        assert st == ref["status"] == 0
        assert np.array_equal(d1, ref["D1"]) and np.array_equal(d2, ref["D2"])
    el.close()


def test_subsampling_refused_by_set_params(pkg):
    el = pkg.Elas(pkg.default_params())
    with pytest.raises(pkg.ElasError) as e:
        el.set_params(pkg.default_params(subsampling=True))
    assert e.value.code == 6
    el.close()
