"""Parity of the HIP path with the oracle, through the C ABI.

Every intermediate of the reference pipeline (descriptors, support lattice
after each pass and filter, support points, disparity grids, Delaunay
meshes, plane prior, dense disparities, final maps) must be bit-exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STAGE_KEYS = ["desc0", "desc1", "can_left", "can_lr", "can_incon", "can_final", "sp0", "sp1",
              "grid0", "grid1", "tri0", "tri1", "mean0", "mean1", "d1_raw", "d2_raw"]


def _compare_all(el, ref, label):
    for k in STAGE_KEYS:
        got = el.fetch(k)
        exp = ref[k]
        assert got.shape == exp.shape, (label, k, got.shape, exp.shape)
        if not np.array_equal(got, exp):
            diff = np.argwhere(got != exp)
            raise AssertionError("%s: stage %s differs at %d positions, first %s (gpu %s, oracle %s)" % (
                label, k, len(diff), diff[0].tolist(), got[tuple(diff[0])], exp[tuple(diff[0])]))


@pytest.mark.parametrize("name", ["450x375_d64", "1242x375_d128"])
def test_every_stage_bit_exact(pkg, oracle, pairs, name):
    left, right, _, dmax = pairs(name)
    print("case", name, left.shape, "dmax", dmax, flush=True)
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    el.set_debug(True)
    D1, D2 = el.process(left, right)
    ref = oracle.process(left, right, oracle.params(disp_max=dmax))
    assert el.status == ref["status"] == 0
    info = el.shape_info()
    assert info["num_support"] == ref["nsp"]
    assert (info["num_tri0"], info["num_tri1"]) == (ref["ntri0"], ref["ntri1"])
    _compare_all(el, ref, name)
    assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])
    assert np.array_equal(el.fetch("D1"), D1)


def test_large_2816x1880_bit_exact(pkg, oracle, pairs):
    """2816x1880, disparities 0..255: final maps and key intermediates."""
    left, right, gt, dmax = pairs("2816x1880_d256")
    print("case 2816x1880", left.shape, flush=True)
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    el.set_debug(True)
    D1, D2 = el.process(left, right)
    ref = oracle.process(left, right, oracle.params(disp_max=dmax))
    _compare_all(el, ref, "2816x1880_d256")
    assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])
    v = D1 >= 0
    assert v.mean() > 0.6 and (np.abs(D1[v] - gt[v]) <= 1).mean() > 0.9


def test_repeat_runs_bitwise_equal(pkg, pairs):
    left, right, _, dmax = pairs("1242x375_d128")
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    outs = [el.process(left, right) for _ in range(3)]
    for D1, D2 in outs[1:]:
        assert np.array_equal(D1, outs[0][0]) and np.array_equal(D2, outs[0][1])


def test_device_resident_equals_host(pkg, pairs):
    import torch
    left, right, _, dmax = pairs("450x375_d64")
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    D1, D2 = el.process(left, right)
    l = torch.from_numpy(left).cuda()
    r = torch.from_numpy(right).cuda()
    d1 = torch.empty(left.shape, dtype=torch.float32, device="cuda")
    d2 = torch.empty_like(d1)
    torch.cuda.synchronize()
    el.process_device(l, r, d1, d2)
    assert np.array_equal(d1.cpu().numpy(), D1) and np.array_equal(d2.cpu().numpy(), D2)


def test_batch_equals_plain(pkg, synth):
    shapes = [(160, 120, 31), (200, 90, 63), (97, 131, 15), (240, 180, 63)]
    pairs_ = []
    for i, (W, H, dm) in enumerate(shapes):
        l, r, _ = synth.make_pair(W, H, dm, regions=4, seed=100 + i)
        pairs_.append((l, r, dm))
    el = pkg.Elas(pkg.default_params(disp_max=63))
    batch = el.process_batch(pairs_)
    for (l, r, dm), (b1, b2, st) in zip(pairs_, batch):
        one = pkg.Elas(pkg.default_params(disp_max=dm))
        D1, D2 = one.process(l, r)
        assert st == one.status
        assert np.array_equal(b1, D1) and np.array_equal(b2, D2)


@pytest.mark.parametrize("slots,chunk", [(1, 2), (3, 1), (3, 2), (2, 3)])
def test_pipelined_batch_equals_plain(pkg, synth, slots, chunk):
    """Chunks moving through pipeline slots give each pair its own result,
    including a textureless pair (status 4) in the middle of the batch."""
    shapes = [(160, 120, 31), (200, 90, 63), (97, 131, 15), (240, 180, 63), (128, 96, 31), (180, 140, 47)]
    pairs_ = []
    for i, (W, H, dm) in enumerate(shapes):
        l, r, _ = synth.make_pair(W, H, dm, regions=4, seed=200 + i)
        pairs_.append((l, r, dm))
    flat = np.full((100, 140), 60, np.uint8)
    pairs_.insert(3, (flat, flat, 31))
    print("pipeline slots", slots, "chunk", chunk, flush=True)
    el = pkg.Elas(pkg.default_params(disp_max=63))
    el.set_pipeline(slots, chunk)
    for rep in range(2):
        batch = el.process_batch(pairs_)
        for (l, r, dm), (b1, b2, st) in zip(pairs_, batch):
            one = pkg.Elas(pkg.default_params(disp_max=dm))
            D1, D2 = one.process(l, r)
            assert st == one.status
            assert np.array_equal(b1, D1) and np.array_equal(b2, D2)


def test_device_batch_equals_plain(pkg, synth):
    import torch
    shapes = [(160, 120, 31), (200, 90, 63), (97, 131, 15), (240, 180, 63), (128, 96, 31)]
    pairs_ = [synth.make_pair(W, H, dm, regions=4, seed=300 + i)[:2] + (dm,) for i, (W, H, dm) in enumerate(shapes)]
    el = pkg.Elas(pkg.default_params(disp_max=63))
    el.set_pipeline(3, 2)
    dev = torch.device("cuda:0")
    L = [torch.from_numpy(l).to(dev) for l, _, _ in pairs_]
    R = [torch.from_numpy(r).to(dev) for _, r, _ in pairs_]
    D1 = [torch.empty(l.shape, dtype=torch.float32, device=dev) for l in L]
    D2 = [torch.empty(l.shape, dtype=torch.float32, device=dev) for l in L]
    st = el.process_batch_device(L, R, D1, D2, [dm for _, _, dm in pairs_])
    for i, (l, r, dm) in enumerate(pairs_):
        one = pkg.Elas(pkg.default_params(disp_max=dm))
        a1, a2 = one.process(l, r)
        assert st[i] == one.status
        assert np.array_equal(D1[i].cpu().numpy(), a1) and np.array_equal(D2[i].cpu().numpy(), a2)


def test_graph_replays_equal_eager(pkg, synth):
    """Phases captured as graphs per shape bucket and replayed with new data,
    interleaved across buckets and with eager runs, equal the eager launches."""
    shapes = [(160, 120), (200, 90), (97, 131)]
    pairs_ = [synth.make_pair(W, H, 63, regions=4, seed=600 + i)[:2] for i, (W, H) in enumerate(shapes * 3)]
    eager = pkg.Elas(pkg.default_params(disp_max=63))
    eager.set_graphs(False)
    graph = pkg.Elas(pkg.default_params(disp_max=63))
    for rep in range(3):
        for i, (l, r) in enumerate(pairs_):
            print("graph replay", rep, i, l.shape, flush=True)
            e1, e2 = eager.process(l, r)
            g1, g2 = graph.process(l, r)
            assert graph.status == eager.status
            assert np.array_equal(e1, g1) and np.array_equal(e2, g2)
    # a parameter change re-captures
    kw = dict(disp_max=63, support_threshold=0.9, lr_threshold=1)
    eager.set_params(pkg.default_params(**kw))
    graph.set_params(pkg.default_params(**kw))
    for l, r in pairs_[:3]:
        e1, e2 = eager.process(l, r)
        g1, g2 = graph.process(l, r)
        assert np.array_equal(e1, g1) and np.array_equal(e2, g2)


@pytest.mark.parametrize("name", ["450x375_d64", "1242x375_d128"])
def test_raster_list_overflow_path(pkg, oracle, pairs, name):
    """Per-tile triangle lists capped at one entry: every tile with triangles
    from two or more chunks takes the overflow path (box tests over the whole
    mesh) and the priors, and everything after them, stay the oracle's."""
    left, right, _, dmax = pairs(name)
    ref = oracle.process(left, right, oracle.params(disp_max=dmax))
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    el.set_debug(True, tiny_lists=True)
    D1, D2 = el.process(left, right)
    _compare_all(el, ref, name + " tiny lists")
    assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])


def test_phase_b_gate_equals_ungated(pkg, pairs, synth):
    """Single-pair calls queue phase B behind a gate while the host
    triangulates; with the gate off phase B is launched afterwards.  Both give
    the same maps, call after call and across shapes."""
    shapes = [pairs("1242x375_d128")[:2] + (pairs("1242x375_d128")[3],)]
    shapes += [synth.make_pair(W, H, dm, regions=3, seed=900 + i)[:2] + (dm,)
               for i, (W, H, dm) in enumerate([(320, 240, 63), (200, 150, 31)])]
    on = pkg.Elas(pkg.default_params(disp_max=127))
    off = pkg.Elas(pkg.default_params(disp_max=127))
    off.set_debug(False, gate=False)
    for rep in range(3):
        for l, r, dm in shapes:
            on.set_params(pkg.default_params(disp_max=dm))
            off.set_params(pkg.default_params(disp_max=dm))
            a1, a2 = on.process(l, r)
            b1, b2 = off.process(l, r)
            assert on.status == off.status == 0
            assert np.array_equal(a1, b1) and np.array_equal(a2, b2), (rep, l.shape, dm)


def test_phase_b_gate_timeout_is_reported(pkg, oracle, pairs):
    """The gate's bounded wait, forced short from the host: a call whose host
    step outlasts it returns an error instead of hanging, and the next call
    without the short bound is the oracle's again."""
    left, right, _, dmax = pairs("450x375_d64")
    ref = oracle.process(left, right, oracle.params(disp_max=dmax), dump=False)
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    D1, D2 = el.process(left, right)
    assert np.array_equal(D1, ref["D1"])
    el.set_debug(False, short_gate=True)
    with pytest.raises(pkg.ElasError, match="gate timed out"):
        el.process(left, right)
    el.set_debug(False)
    for _ in range(3):
        D1, D2 = el.process(left, right)
        assert el.status == 0
        assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])


def test_sparse_texture_large_triangles(pkg, oracle, synth):
    """Texture only in five patches: twenty triangles per mesh, most spanning
    a large part of the image, so the raster's waves walk long column ranges
    and many triangles cover each pixel's neighbourhood."""
    l, r, _ = synth.make_pair(240, 180, 31, regions=3, seed=808)
    mask = np.zeros(l.shape, bool)
    for y, x in [(20, 30), (30, 180), (120, 40), (130, 170), (80, 110)]:
        mask[y:y + 24, x:x + 24] = True
    left = np.where(mask, l, 100).astype(np.uint8)
    right = np.where(mask, r, 100).astype(np.uint8)
    ref = oracle.process(left, right, oracle.params(disp_max=31))
    assert ref["status"] == 0 and ref["ntri0"] >= 10
    el = pkg.Elas(pkg.default_params(disp_max=31))
    el.set_debug(True)
    D1, D2 = el.process(left, right)
    _compare_all(el, ref, "sparse texture")
    assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])


def test_textureless_pair_reports_too_few_points(pkg, oracle):
    img = np.full((120, 160), 77, np.uint8)
    el = pkg.Elas(pkg.default_params(disp_max=31))
    D1, D2 = el.process(img, img)
    assert el.status == pkg.ELAS_ERR_TOO_FEW_POINTS
    assert (D1 == -1).all() and (D2 == -1).all()
    ref = oracle.process(img, img, oracle.params(disp_max=31))
    assert ref["status"] == pkg.ELAS_ERR_TOO_FEW_POINTS


@pytest.mark.parametrize("W,H", [(1, 1), (11, 9), (37, 23), (64, 64)])
def test_tiny_images(pkg, oracle, synth, W, H):
    l, r, _ = synth.make_pair(W, H, 15, regions=2, seed=W * 1000 + H) if W > 8 and H > 8 else (
        np.zeros((H, W), np.uint8), np.zeros((H, W), np.uint8), None)
    el = pkg.Elas(pkg.default_params(disp_max=15))
    D1, D2 = el.process(l, r)
    ref = oracle.process(l, r, oracle.params(disp_max=15), dump=False)
    assert el.status == ref["status"]
    assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])


def test_non_default_parameters(pkg, oracle, synth):
    """Every parameter the kernels read is honoured (the reference hardwires macros)."""
    l, r, _ = synth.make_pair(300, 200, 47, regions=5, seed=11)
    kw = dict(disp_max=47, support_threshold=0.9, support_texture=6, candidate_stepsize=4,
              incon_window_size=4, incon_threshold=4, incon_min_support=4, grid_size=16,
              gamma=4.0, sradius=3.0, lr_threshold=1, disp_min=2)
    el = pkg.Elas(pkg.default_params(**kw))
    el.set_debug(True)
    D1, D2 = el.process(l, r)
    ref = oracle.process(l, r, oracle.params(**kw))
    _compare_all(el, ref, "params")
    assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])
