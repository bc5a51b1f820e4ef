"""Post-processing filters on the GPU (speckle removal, gap interpolation,
adaptive mean, median) against the oracle's CPU definitions: bit-exact final
maps.  The reference declares these filters without implementing them, so the
oracle's definitions (pinned in test_oracle_post.py) are the specification."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = {
    "speckle": dict(speckle_size=200, speckle_sim_threshold=1.0),
    "speckle_small_sim": dict(speckle_size=40, speckle_sim_threshold=0.0),
    "gap": dict(ipol_gap_width=5000),
    "gap_short": dict(ipol_gap_width=7),
    "adaptive": dict(filter_adaptive_mean=True),
    "median": dict(filter_median=True),
    "all": dict(speckle_size=200, speckle_sim_threshold=1.0, ipol_gap_width=5000, filter_adaptive_mean=True,
                filter_median=True),
    "all_left_only": dict(speckle_size=100, speckle_sim_threshold=2.0, ipol_gap_width=300, filter_median=True,
                          postprocess_only_left=True),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_post_filters_bit_exact(pkg, oracle, synth, case):
    l, r, _ = synth.make_pair(300, 200, 47, regions=6, seed=21, occlusion=0.08)
    kw = dict(disp_max=47, **CASES[case])
    print("post case", case, flush=True)
    el = pkg.Elas(pkg.default_params(**kw))
    D1, D2 = el.process(l, r)
    ref = oracle.process(l, r, oracle.params(**kw), dump=False)
    for name, got, exp in (("D1", D1, ref["D1"]), ("D2", D2, ref["D2"])):
        if not np.array_equal(got, exp):
            bad = np.argwhere(got != exp)
            raise AssertionError("%s %s differs at %d px, first %s (gpu %r, oracle %r)" % (
                case, name, len(bad), bad[0].tolist(), got[tuple(bad[0])], exp[tuple(bad[0])]))


def test_post_filters_1242x375_declared_defaults(pkg, oracle, pairs):
    """The Elas.h initialisers (post-processing on) on the 1242x375 workload."""
    left, right, _, dmax = pairs("1242x375_d128")
    dp = pkg.declared_params(disp_max=dmax)
    el = pkg.Elas(dp)
    D1, D2 = el.process(left, right)
    prm = oracle.params(**{k: getattr(dp, k) for k, _ in dp._fields_})
    ref = oracle.process(left, right, prm, dump=False)
    assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])


def test_post_filters_in_a_batch(pkg, synth):
    kw = CASES["all"]
    shapes = [(160, 120, 31), (97, 131, 15), (240, 180, 63)]
    pairs_ = [synth.make_pair(W, H, dm, regions=4, seed=40 + i)[:2] + (dm,) for i, (W, H, dm) in enumerate(shapes)]
    el = pkg.Elas(pkg.default_params(disp_max=63, **kw))
    el.set_pipeline(2, 1)
    out = el.process_batch(pairs_)
    for (l, r, dm), (b1, b2, st) in zip(pairs_, out):
        one = pkg.Elas(pkg.default_params(disp_max=dm, **kw))
        D1, D2 = one.process(l, r)
        assert np.array_equal(b1, D1) and np.array_equal(b2, D2)
