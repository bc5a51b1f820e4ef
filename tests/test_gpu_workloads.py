"""Oracle parity of the batched and served workloads at their real shapes, and
of pairs beyond the reference's envelope, through the C ABI.

The mix is 64 pairs drawn 3:2:1 from 640x480 d0..63, 1242x375 d0..127 and
1920x1080 d0..255 (pair i seeded 3000 + i).  Each batched or served pair must
give the maps Elas::process gives for it alone (src/Elas.cu:19-120), which
the oracle restates: bit-exact D1/D2 for every pair.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STAGE_KEYS = ["desc0", "desc1", "can_left", "can_lr", "can_incon", "can_final", "sp0", "sp1",
              "grid0", "grid1", "tri0", "tri1", "mean0", "mean1", "d1_raw", "d2_raw"]


@pytest.fixture(scope="module")
def mix(synth, oracle):
    pairs = synth.batch_pairs(64)
    print("oracle over the 64-pair mix", flush=True)
    refs = oracle.process_many(pairs)
    assert all(st == 0 for _, _, st in refs)
    return pairs, refs


def _same_maps(label, got1, got2, ref):
    for name, g, e in (("D1", got1, ref[0]), ("D2", got2, ref[1])):
        if not np.array_equal(g, e):
            diff = np.argwhere(g != e)
            raise AssertionError("%s %s: %d pixels differ, first %s (gpu %s, oracle %s)" % (
                label, name, len(diff), diff[0].tolist(), g[tuple(diff[0])], e[tuple(diff[0])]))


def _stages_equal(pkg, oracle, left, right, dmax, label):
    el = pkg.Elas(pkg.default_params(disp_max=dmax))
    el.set_debug(True)
    D1, D2 = el.process(left, right)
    ref = oracle.process(left, right, oracle.params(disp_max=dmax))
    assert el.status == ref["status"] == 0, (label, el.status, ref["status"])
    for k in STAGE_KEYS:
        got, exp = el.fetch(k), ref[k]
        assert got.shape == exp.shape, (label, k)
        if not np.array_equal(got, exp):
            diff = np.argwhere(got != exp)
            raise AssertionError("%s: stage %s differs at %d positions, first %s" % (label, k, len(diff),
                                                                                    diff[0].tolist()))
    _same_maps(label, D1, D2, (ref["D1"], ref["D2"]))
    el.close()


def test_mix_shapes_every_stage_bit_exact(pkg, oracle, synth):
    """The first pair of each shape of the mix, every intermediate."""
    seen = set()
    for i, (l, r, dm) in enumerate(synth.batch_pairs(64)):
        key = (l.shape, dm)
        if key in seen:
            continue
        seen.add(key)
        print("mix pair", i, l.shape, "dmax", dm, flush=True)
        _stages_equal(pkg, oracle, l, r, dm, "mix pair %d %dx%d d%d" % (i, l.shape[1], l.shape[0], dm + 1))
    assert len(seen) == 3


@pytest.mark.parametrize("slots,chunk", [(4, 8), (3, 16)])
def test_mix_batch_device_bit_exact(pkg, mix, slots, chunk):
    import torch
    pairs, refs = mix
    print("device batch of 64, slots", slots, "chunk", chunk, flush=True)
    dev = torch.device("cuda:0")
    L = [torch.from_numpy(l).to(dev) for l, _, _ in pairs]
    R = [torch.from_numpy(r).to(dev) for _, r, _ in pairs]
    D1 = [torch.empty(l.shape, dtype=torch.float32, device=dev) for l in L]
    D2 = [torch.empty(l.shape, dtype=torch.float32, device=dev) for l in L]
    el = pkg.Elas(pkg.default_params(disp_max=255))
    el.set_pipeline(slots, chunk)
    for rep in range(2):  # the second run replays the captured graphs
        for d in D1 + D2:
            d.fill_(-7.0)
        st = el.process_batch_device(L, R, D1, D2, [dm for _, _, dm in pairs])
        assert st == [0] * len(pairs)
        for i, ref in enumerate(refs):
            _same_maps("rep %d pair %d" % (rep, i), D1[i].cpu().numpy(), D2[i].cpu().numpy(), ref)
    el.close()


def test_mix_batch_host_buffers_bit_exact(pkg, mix):
    pairs, refs = mix
    print("host batch of 64", flush=True)
    el = pkg.Elas(pkg.default_params(disp_max=255))
    el.set_pipeline(3, 16)
    out = el.process_batch(pairs)
    for i, ((d1, d2, st), ref) in enumerate(zip(out, refs)):
        assert st == 0
        _same_maps("pair %d" % i, d1, d2, ref)
    el.close()


def test_mix_served_poisson_bit_exact(pkg, mix):
    """Requests of the mix arrive as a Poisson stream (400 pairs/s, seed 4000)
    at the continuous-batching loop; every served result equals the oracle's
    for that pair alone."""
    pairs, refs = mix
    n = 40
    idx = list(range(n))
    shapes = {(pairs[i][0].shape, pairs[i][2]) for i in idx}
    assert len(shapes) == 3
    rng = np.random.default_rng(4000)
    arrive = np.cumsum(rng.exponential(1.0 / 400.0, n))
    el = pkg.Elas(pkg.default_params(disp_max=255))
    el.serve_start(hbm_budget_bytes=4 << 30, max_chunk_pairs=8, slots=3)
    outs = [(np.empty(pairs[i][0].shape, np.float32), np.empty(pairs[i][0].shape, np.float32)) for i in idx]
    print("serving", n, "requests of the mix at 400 pairs/s", flush=True)
    t0 = time.perf_counter()
    tickets = []
    for k, i in enumerate(idx):
        wait = t0 + arrive[k] - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        l, r, dm = pairs[i]
        tickets.append(el.submit(l, r, outs[k][0], outs[k][1], dm))
    done = {}
    while len(done) < n:
        for t in tickets:
            if t not in done:
                res = el.poll(t)
                if res is not None:
                    done[t] = res
        assert time.perf_counter() - t0 < 90, "serving loop did not finish"
        time.sleep(0.0005)
    stats = el.serve_stats()
    el.serve_stop()
    for k, (i, t) in enumerate(zip(idx, tickets)):
        assert done[t][0] == 0, (i, done[t])
        _same_maps("request %d (pair %d)" % (k, i), outs[k][0], outs[k][1], refs[i])
    assert stats["completed"] == n and stats["chunks"] >= 2
    el.close()


@pytest.mark.parametrize("W,H,dmax", [(4096, 256, 255), (5000, 120, 511)])
def test_beyond_the_reference_envelope(pkg, oracle, synth, W, H, dmax):
    """Wider than the reference's 3072-column shared-memory rows
    (src/Elas.cu:506-508) and with 512 disparities in a single-pair call:
    every intermediate bit-exact."""
    l, r, _ = synth.make_pair(W, H, dmax, regions=6, seed=W + dmax, slope=0.05)
    print("case %dx%d dmax %d" % (W, H, dmax), flush=True)
    _stages_equal(pkg, oracle, l, r, dmax, "%dx%d d%d" % (W, H, dmax + 1))
