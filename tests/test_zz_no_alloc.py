"""The entry point that allocates nothing (elas_hip_process_device_noalloc):
its maps equal elas_hip_process_device's, bit for bit, on two calls in a row
(the second with a smaller pair, so the scratch still holds the first call's
larger contents), and a pair that does not fit the scratch is refused before
any work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(pkg, l, r):
    import torch
    dev = torch.device("cuda:0")
    L = torch.from_numpy(l).to(dev)
    R = torch.from_numpy(r).to(dev)
    D1 = torch.full(l.shape, -7.0, dtype=torch.float32, device=dev)
    D2 = torch.full(l.shape, -7.0, dtype=torch.float32, device=dev)
    return L, R, D1, D2


def test_no_alloc_equals_process_device(pkg, synth):
    cases = [(1242, 375, 127, 1), (640, 480, 63, 2)]   # the second call is smaller
    ref = pkg.Elas(pkg.default_params(disp_max=127))
    el = pkg.Elas(pkg.default_params(disp_max=127))
    for W, H, dmax, seed in cases:
        print("case %dx%d dmax %d" % (W, H, dmax), flush=True)
        l, r, _ = synth.make_pair(W, H, dmax, regions=8, seed=seed, slope=0.05)
        ref.set_params(pkg.default_params(disp_max=dmax))
        el.set_params(pkg.default_params(disp_max=dmax))
        L, R, A1, A2 = _dev(pkg, l, r)
        _, _, B1, B2 = _dev(pkg, l, r)
        assert ref.process_device(L, R, A1, A2) == 0
        assert el.process_device_noalloc(L, R, B1, B2) == 0
        a1, a2, b1, b2 = (t.cpu().numpy() for t in (A1, A2, B1, B2))
        assert (a1 >= 0).mean() > 0.5
        assert np.array_equal(a1, b1) and np.array_equal(a2, b2), (W, H, int((a1 != b1).sum()))
    # a repeat of the first shape after the smaller one, on the same scratch
    l, r, _ = synth.make_pair(1242, 375, 127, regions=8, seed=1, slope=0.05)
    ref.set_params(pkg.default_params(disp_max=127))
    el.set_params(pkg.default_params(disp_max=127))
    L, R, A1, A2 = _dev(pkg, l, r)
    _, _, B1, B2 = _dev(pkg, l, r)
    ref.process_device(L, R, A1, A2)
    el.process_device_noalloc(L, R, B1, B2)
    assert np.array_equal(A1.cpu().numpy(), B1.cpu().numpy()) and np.array_equal(A2.cpu().numpy(), B2.cpu().numpy())
    ref.close()
    el.close()


def test_no_alloc_refuses_a_pair_larger_than_the_scratch(pkg):
    import torch
    el = pkg.Elas(pkg.default_params(disp_max=255))
    W, H = 2816, 1880
    L = torch.zeros((H, W), dtype=torch.uint8, device="cuda:0")
    D = torch.full((H, W), -7.0, dtype=torch.float32, device="cuda:0")
    with pytest.raises(pkg.ElasError) as e:
        el.process_device_noalloc(L, L, D, D)
    assert e.value.code == 6 and "scratch" in el.last_error()
    assert bool((D == -7.0).all())   # refused before any work
    el.close()
