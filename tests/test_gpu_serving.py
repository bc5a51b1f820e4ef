"""Serving loop on the GPU: continuous batching over pipeline slots.

Every result must be bit-identical to the same pair processed alone by a
plain Elas::process call, whatever else was in flight with it."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mix(synth, n, seed0):
    shapes = [(160, 120, 31), (200, 90, 63), (97, 131, 15), (240, 180, 63), (128, 96, 31), (300, 200, 47)]
    out = []
    for i in range(n):
        W, H, dm = shapes[i % len(shapes)]
        l, r, _ = synth.make_pair(W, H, dm, regions=4, seed=seed0 + i)
        out.append((l, r, dm))
    return out


def _alone(pkg, pairs):
    res = []
    for l, r, dm in pairs:
        e = pkg.Elas(pkg.default_params(disp_max=dm))
        D1, D2 = e.process(l, r)
        res.append((D1, D2, e.status))
        e.close()
    return res


def _drain(el, tickets, timeout=60.0):
    done = {}
    t0 = time.time()
    while len(done) < len(tickets):
        for t in tickets:
            if t not in done:
                r = el.poll(t)
                if r is not None:
                    done[t] = r
        assert time.time() - t0 < timeout, "serving loop did not finish"
        time.sleep(0.0005)
    return done


def test_results_same_alone_and_in_the_mix(pkg, synth):
    pairs = _mix(synth, 18, 500)
    flat = np.full((90, 130), 200, np.uint8)
    pairs.insert(7, (flat, flat.copy(), 31))       # too few support points, in the middle
    ref = _alone(pkg, pairs)
    el = pkg.Elas(pkg.default_params(disp_max=63))
    el.serve_start(hbm_budget_bytes=256 << 20, max_chunk_pairs=4, slots=3)
    outs = [(np.empty(l.shape, np.float32), np.empty(l.shape, np.float32)) for l, _, _ in pairs]
    tickets = []
    print("serving", len(pairs), "pairs in bursts", flush=True)
    for i, (l, r, dm) in enumerate(pairs):           # staggered bursts: later pairs join chunks in flight
        tickets.append(el.submit(l, r, outs[i][0], outs[i][1], dm))
        if i % 5 == 4:
            time.sleep(0.002)
    done = _drain(el, tickets)
    stats = el.serve_stats()
    el.serve_stop()
    for i, t in enumerate(tickets):
        st, lat = done[t]
        assert st == ref[i][2], (i, st, ref[i][2])
        assert np.array_equal(outs[i][0], ref[i][0]) and np.array_equal(outs[i][1], ref[i][1]), i
        assert lat > 0
    assert stats["completed"] == len(pairs) and stats["submitted"] == len(pairs)
    assert stats["max_chunk_pairs"] <= 4
    assert stats["peak_in_flight_bytes"] <= 256 << 20
    el.close()


def test_device_resident_requests(pkg, synth):
    import torch
    pairs = _mix(synth, 6, 700)
    ref = _alone(pkg, pairs)
    dev = torch.device("cuda:0")
    el = pkg.Elas(pkg.default_params(disp_max=63))
    el.serve_start(hbm_budget_bytes=128 << 20, max_chunk_pairs=3)
    bufs = []
    tickets = []
    for l, r, dm in pairs:
        L, R = torch.from_numpy(l).to(dev), torch.from_numpy(r).to(dev)
        D1 = torch.empty(l.shape, dtype=torch.float32, device=dev)
        D2 = torch.empty_like(D1)
        bufs.append((D1, D2))
        tickets.append(el.submit(L, R, D1, D2, dm))
    done = _drain(el, tickets)
    el.serve_stop()
    for i, t in enumerate(tickets):
        assert done[t][0] == ref[i][2]
        assert np.array_equal(bufs[i][0].cpu().numpy(), ref[i][0])
        assert np.array_equal(bufs[i][1].cpu().numpy(), ref[i][1])
    el.close()


def test_admission_limits_and_busy(pkg, synth):
    el = pkg.Elas(pkg.default_params(disp_max=63))
    el.serve_start(hbm_budget_bytes=3 * (8 << 20), max_chunk_pairs=8, slots=3)   # 8 MiB per slot
    big = np.zeros((600, 800), np.uint8)
    with pytest.raises(pkg.ElasError) as e:                                        # needs > 8 MiB
        el.submit(big, big, np.empty((600, 800), np.float32), np.empty((600, 800), np.float32), 63)
    assert e.value.code == 3
    with pytest.raises(pkg.ElasError) as e:                                        # blocking calls refused
        el.process(big, big)
    assert e.value.code == 7
    pairs = _mix(synth, 4, 900)
    outs = [(np.empty(l.shape, np.float32), np.empty(l.shape, np.float32)) for l, _, _ in pairs]
    tickets = [el.submit(l, r, o[0], o[1], dm) for (l, r, dm), o in zip(pairs, outs)]
    _drain(el, tickets)
    st = el.serve_stats()
    assert st["peak_in_flight_bytes"] <= 3 * (8 << 20)
    el.serve_stop()
    l, r, _ = pairs[0]
    D1, D2 = el.process(l, r)                                                        # usable again
    ref = _alone(pkg, [(l, r, 63)])[0]
    assert np.array_equal(D1, ref[0]) and np.array_equal(D2, ref[1])
    el.close()
