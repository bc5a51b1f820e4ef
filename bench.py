"""Benchmark of the ELAS path on one MI355X.

Workload (the largest single-GPU configuration of the headline metric): one
synthetic 2816x1880 uint8 stereo pair, disparities 0..255, seed 2 (32
slanted-plane regions, slopes up to 0.1), reference-effective parameters.  A
step is one Elas::process of that pair with images and disparity maps
resident in HBM (elas_hip_process_device), host Delaunay included.

value = Mpx*disp/s = W*H*(disp_max+1)/1e6 per pair / wall seconds over the
timed steps.  Also reported: pairs/s, per-call median and p90 latency, the
same pair through the host-buffer call (elas_hip_process, PCIe included),
per-stage device times (HIP events on the pipeline stream), per-kernel HBM
and SAD fractions with the roofline of the dominant kernel, the
batched-vs-plain pairs/s on the mixed 64-pair batch, the serving sweep, and
the oracle (the reference algorithm restated on the CPU) timed on this host
on one core and on every core.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0
# v_sad_u8 (4 byte-|diff| per lane-op): 256 CU x 64 lanes x 2.4 GHz x 4
SAD_PEAK_TBYTE = 256 * 64 * 2.4e9 * 4 / 1e12


def work_counts(oracle, left, right, dmax):
    """Algorithmic work of the kernels, counted by the oracle on this input:
    (candidate, d) pairs of each support pass and (pixel, d) pairs of dense
    matching."""
    prm = oracle.params(disp_max=dmax)
    d0 = oracle.descriptor(left)
    d1 = oracle.descriptor(right)
    H, W = left.shape
    can = np.full((H // prm.candidate_stepsize, W // prm.candidate_stepsize), -1, np.int32)
    w_left = oracle.support_pass(d0, d1, prm, 0, can)
    w_right = oracle.support_pass(d0, d1, prm, 1, can)
    out = oracle.process(left, right, prm)
    return dict(support_left=w_left, support_right=w_right, dense=out["work_dense"])


def batch_vs_plain(pkg, synth, n=64, reps=3, pool=None):
    """The mixed batch: n pairs drawn 3:2:1 from 640x480/d64, 1242x375/d128,
    1920x1080/d256 (seeds 3000+i), images and disparity maps resident in HBM.
    Plain: n elas_hip_process_device calls, one pair at a time.  Batched: one
    elas_hip_process_batch_device call, pairs cut into chunks that move
    through pipeline slots (one launch per stage per chunk; the Delaunay step
    of one chunk overlaps the GPU stages of the others).  The (slots, chunk)
    setting is picked by timing a few candidates; outputs must be identical
    to the plain ones.  The host-buffer form (PCIe transfers included) is
    timed once with the chosen setting."""
    import torch
    pairs = pool[:n] if pool else synth.batch_pairs(n)
    work = sum(l.shape[0] * l.shape[1] * (dm + 1) for l, _, dm in pairs) / 1e6
    dev = torch.device("cuda:0")
    L = [torch.from_numpy(l).to(dev) for l, _, _ in pairs]
    R = [torch.from_numpy(r).to(dev) for _, r, _ in pairs]
    mk = lambda: [torch.empty(l.shape, dtype=torch.float32, device=dev) for l in L]
    P1, P2, B1, B2 = mk(), mk(), mk(), mk()
    dms = [dm for _, _, dm in pairs]
    ctxs = {}
    for dm in dms:
        if dm not in ctxs:
            ctxs[dm] = pkg.Elas(pkg.default_params(disp_max=dm))
    torch.cuda.synchronize()

    def plain():
        for i in range(n):
            ctxs[dms[i]].process_device(L[i], R[i], P1[i], P2[i])
        torch.cuda.synchronize()

    eb = pkg.Elas(pkg.default_params(disp_max=255))

    def batched():
        eb.process_batch_device(L, R, B1, B2, dms)
        torch.cuda.synchronize()

    def best_of(f):
        f()                                       # warm-up (allocations)
        t = 1e9
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            t = min(t, time.perf_counter() - t0)
        return t

    tp = best_of(plain)
    tune = {}
    for slots, chunk in ((1, 0), (2, 16), (3, 8), (3, 16), (4, 8), (3, 22), (4, 16)):
        eb.set_pipeline(slots, chunk)
        tune["%d/%d" % (slots, chunk)] = best_of(batched)
    pick = min(tune, key=tune.get)
    slots, chunk = (int(x) for x in pick.split("/"))
    eb.set_pipeline(slots, chunk)
    tb = best_of(batched)
    same = all(torch.equal(B1[i], P1[i]) and torch.equal(B2[i], P2[i]) for i in range(n))
    # host buffers in and out (PCIe-inclusive), same pipeline setting
    hp = [(l, r, dm) for l, r, dm in pairs]
    eb.process_batch(hp)
    t0 = time.perf_counter()
    out_h = eb.process_batch(hp)
    th = time.perf_counter() - t0
    same_h = all(np.array_equal(out_h[i][0], P1[i].cpu().numpy()) and np.array_equal(out_h[i][1], P2[i].cpu().numpy())
                 for i in range(n))
    for c in list(ctxs.values()) + [eb]:
        c.close()
    return {"pairs": n, "mpx_disp": round(work, 1), "batched_pairs_per_s": round(n / tb, 2),
            "plain_pairs_per_s": round(n / tp, 2), "batched_mpx_disp_per_s": round(work / tb, 1),
            "plain_mpx_disp_per_s": round(work / tp, 1), "speedup": round(tp / tb, 3),
            "identical_outputs": bool(same and same_h), "batch_ms": round(tb * 1e3, 2),
            "plain_ms": round(tp * 1e3, 2), "pipeline": {"slots": slots, "chunk": chunk},
            "tuning_ms": {k: round(v * 1e3, 2) for k, v in tune.items()},
            "host_buffers_batched_pairs_per_s": round(n / th, 2), "buffers": "device-resident"}


def serving(pkg, synth, rates=(50, 100, 200, 400, 800), n_req=2000, pool=None, seed=4000,
            knee=(1200, 1600, 2400, 3200, 4800, 6400)):
    """Serving sweep: Poisson arrivals at each rate of pairs drawn from the
    mixed-batch pool (request i uses pool pair i mod len(pool)), host buffers
    in pinned memory, submitted to the continuous-batching loop
    (elas_hip_submit / elas_hip_poll).  Latency = arrival -> D1/D2 in host
    memory (the feeder's own lateness included); sustained pairs/s = requests
    / (first arrival -> last completion).  After `rates`, the `knee` rates run
    in turn until one is past the knee: less than 90% of the offered rate
    sustained, or the median latency above twice the lowest rate's (queueing
    builds up; the p99 of the mix is set by its largest pairs at any rate)."""
    import torch
    pool = pool or synth.batch_pairs(64)
    pin = lambda a: torch.from_numpy(a).pin_memory().numpy()
    P = [(pin(l), pin(r), dm, torch.empty(l.shape, dtype=torch.float32, pin_memory=True).numpy(),
          torch.empty(l.shape, dtype=torch.float32, pin_memory=True).numpy()) for l, r, dm in pool]
    el = pkg.Elas(pkg.default_params(disp_max=255))
    # one pair of every shape first: the dense tiling of a new shape is tuned
    # after its first call, which should not land in the latencies
    seen = {}
    for l, r, dm in pool:
        seen.setdefault((l.shape, dm), (l, r, dm))
    el.process_batch(list(seen.values()))
    out = {}
    rng = np.random.default_rng(seed)
    for lam in list(rates) + list(knee):
        if lam in knee and out and out[max(out, key=float)].get("past_knee"):
            break
        el.serve_start()
        gaps = rng.exponential(1.0 / lam, n_req)
        lat, late, pending = [], [], {}
        t0 = time.perf_counter() + 0.01
        arr = t0 + np.cumsum(gaps)
        t_last = t0
        i = 0
        while i < n_req or pending:
            now = time.perf_counter()
            if i < n_req and now >= arr[i]:
                l, r, dm, d1, d2 = P[i % len(P)]
                pending[el.submit(l, r, d1, d2, dm)] = arr[i]
                late.append(time.perf_counter() - arr[i])
                i += 1
                continue
            for t in list(pending)[:64]:   # the oldest requests (completions are nearly in order)
                res = el.poll(t)
                if res is not None:
                    a = pending.pop(t)
                    done = time.perf_counter()
                    t_last = max(t_last, done)
                    lat.append((done - a) * 1e3)
            if i < n_req and not pending:
                time.sleep(max(0.0, min(arr[i] - time.perf_counter() - 2e-4, 0.002)))
        st = el.serve_stats()
        el.serve_stop()
        lat = np.array(lat)
        out[str(lam)] = {"requests": n_req, "p50_ms": round(float(np.percentile(lat, 50)), 3),
                         "p99_ms": round(float(np.percentile(lat, 99)), 3),
                         "mean_ms": round(float(lat.mean()), 3),
                         "sustained_pairs_per_s": round(n_req / (t_last - t0), 1),
                         "feeder_late_p99_ms": round(float(np.percentile(late, 99)) * 1e3, 3),
                         "chunks": int(st["chunks"]), "max_chunk_pairs": int(st["max_chunk_pairs"]),
                         "max_in_flight_pairs": int(st["max_in_flight_pairs"]),
                         "peak_in_flight_MB": round(st["peak_in_flight_bytes"] / 1e6, 1)}
        r = out[str(lam)]
        p50_low = min(v["p50_ms"] for v in out.values())
        r["past_knee"] = bool(r["sustained_pairs_per_s"] < 0.9 * lam or r["p50_ms"] > 2 * p50_low)
    el.close()
    return {"rates": out, "pool_pairs": len(P), "hbm_budget_GiB": 8, "buffers": "pinned host"}


def kernel_rooflines(el, wc, W, H, dmax, reps=50, warm_ms=100.0):
    """Per-kernel roofline of the last call.  Each kernel's average launch
    time comes from HIP events around `reps` back-to-back re-launches on the
    pipeline stream (elas_hip_time_kernel), after about 100 ms of untimed
    re-launches (cold_ms: the first 10 launches).  Algorithmic work per launch:
      descriptor  17 B per pixel and image (1 B image read, 16 B written)
      support     64 byte-|diffs| per evaluated (candidate, d) (oracle count);
                  HBM: descriptor rows v-2 and v+2 of both images once per
                  lattice row (64 B per lattice-row pixel, both passes), and
                  the lattice written by both passes (8 B per candidate)
      dense       16 byte-|diffs| per evaluated (pixel, d) (oracle count);
                  HBM: both descriptor images read once, both priors read,
                  both dense maps written: 48 B per pixel of the pair
      lr          16 B per pixel (read D1, D2; write D1, D2)
    The bound is the one with the longer time at its peak (HBM 8 TB/s,
    v_sad_u8 157.3 T byte-|diff|/s)."""
    cw, ch = W // 5, H // 5
    work = {
        "descriptor": (0.0, 2 * 17.0 * W * H),
        # both passes in one launch need rows v-2 and v+2 of both descriptor
        # images at each lattice row v (2 rows x 2 images x 16 B per pixel)
        "support": (64.0 * (wc["support_left"] + wc["support_right"]), 64.0 * W * ch + 8.0 * cw * ch),
        "dense": (16.0 * wc["dense"], 48.0 * W * H),
        "lr": (0.0, 16.0 * W * H),
    }
    out = {}
    for k, (ops, byts) in work.items():
        # the first launches after the host-bound steps run while the GPU's
        # clock is still rising; the kernel is launched back to back for about
        # 100 ms first, so the figure is its steady-state time
        cold = el.time_kernel(k, 10)
        if warm_ms > 0:
            el.time_kernel(k, max(10, min(5000, int(warm_ms / max(cold, 1e-3)))))
        ms = el.time_kernel(k, reps)
        t_sad = ops / (SAD_PEAK_TBYTE * 1e12)
        t_hbm = byts / (HBM_PEAK_GBS * 1e9)
        if t_sad >= t_hbm:
            out[k] = dict(bound="valu_sad", achieved=ops / (ms * 1e-3) / 1e12, peak=SAD_PEAK_TBYTE,
                          unit="Tbyte-diff/s", ms=ms, algorithmic=ops)
        else:
            out[k] = dict(bound="hbm", achieved=byts / (ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                          ms=ms, algorithmic=byts)
        out[k]["frac"] = out[k]["achieved"] / out[k]["peak"]
        out[k]["cold_ms"] = cold
        out[k]["sad_frac"] = (ops / (ms * 1e-3) / 1e12) / SAD_PEAK_TBYTE
        out[k]["hbm_frac"] = (byts / (ms * 1e-3) / 1e9) / HBM_PEAK_GBS
    return out


KERNEL_SYMBOL = {"descriptor": "k_descriptor", "support": "k_support_lr", "dense": "k_dense<", "lr": "k_lr"}


def traffic_of(kernel, workload, W, H, dmax):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    profiles/hbm_traffic_<W>x<H>_d<dmax+1>.txt (FETCH_SIZE x2 +
    WRITE_SIZE x1, scaled by a plain device copy), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%dx%d_d%d.txt" % (W, H, dmax + 1))
    if not os.path.isfile(path):
        return None
    for line in open(path):
        parts = line.split()
        sym = KERNEL_SYMBOL.get(kernel, "?")
        if len(parts) >= 4 and (parts[0] == sym or (sym.endswith("<") and parts[0].startswith(sym))):
            try:
                return int(float(parts[3]))
            except ValueError:
                return None
    return None


def cpu_baseline(oracle, left, right, dmax, budget_s=12.0):
    """The oracle (the reference algorithm restated in C++, one thread per
    pipeline) on this host: whole pipelines of the same pair back to back on
    one core for about budget_s, then one pipeline per logical CPU at once
    (processes, since Triangle keeps its state in globals), timed after a
    warm-up round of the same pool."""
    import concurrent.futures as cf
    import multiprocessing as mp
    prm = oracle.params(disp_max=dmax)
    n, t0 = 0, time.perf_counter()
    while True:
        oracle.process(left, right, prm, dump=False)
        n += 1
        el = time.perf_counter() - t0
        if (el > budget_s and n >= 2) or n >= 1000:
            break
    cores = os.cpu_count() or 1
    with cf.ProcessPoolExecutor(cores, mp_context=mp.get_context("spawn")) as ex:
        list(ex.map(oracle._maps_job, [(left[:64, :64].copy(), right[:64, :64].copy(), 15)] * cores))
        t1 = time.perf_counter()
        list(ex.map(oracle._maps_job, [(left, right, dmax)] * cores))
        el_all = time.perf_counter() - t1
    return n, el, cores, el_all


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched comparison")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving sweep")
    ap.add_argument("--serve-quick", action="store_true",
                    help="reduced serving sweep (100, 400, 800 pairs/s, 400 requests each, no knee search)")
    ap.add_argument("--workload", default="2816x1880_d256", choices=["450x375_d64", "1242x375_d128", "2816x1880_d256"],
                    help="2816x1880_d256 is the headline; the others for profiling")
    ap.add_argument("--tri-split", type=int, default=-1,
                    help="host Delaunay split depth (0 serial Triangle, -1 automatic)")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer single-pair rate")
    ap.add_argument("--dense-tiles", default="",
                    help="force the dense tiling 'cells_per_wave,steps_per_chunk' (profiler runs, where the "
                         "autotuner's timings are distorted)")
    ap.add_argument("--kernel-warm-ms", type=float, default=100.0,
                    help="back-to-back launches of each kernel before its timed ones (0: none; for counter runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy a 256 MiB device buffer 5 times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (pairs would shard across GPUs with no "
              "exchange; no multi-GPU path is built)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        src.fill_(1)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 256 << 20, "copies": 5}))
        return
    pkg = __graft_entry__.load_package()
    synth = __graft_entry__.load_synth()
    import oracle

    left, right, _, dmax = synth.config_pair(args.workload)
    H, W = left.shape
    mpxd = W * H * (dmax + 1) / 1e6
    pkg.set_tri_split(args.tri_split)
    el = pkg.Elas(pkg.default_params(disp_max=dmax), device=0)
    if args.dense_tiles:
        el.set_dense_tiles(*(int(x) for x in args.dense_tiles.split(",")))
    dev = torch.device("cuda:0")
    L = torch.from_numpy(left).to(dev)
    R = torch.from_numpy(right).to(dev)
    D1 = torch.empty((H, W), dtype=torch.float32, device=dev)
    D2 = torch.empty_like(D1)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        el.process_device(L, R, D1, D2)
    torch.cuda.synchronize()
    # each call is synchronous (it returns with the maps in HBM), so the
    # per-call clock brackets one whole step
    per_call = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tc = time.perf_counter()
        el.process_device(L, R, D1, D2)
        per_call.append(time.perf_counter() - tc)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_call = np.array(per_call) * 1e3
    # per-stage timers of the same calls, read in separate untimed steps so
    # that reading them stays out of the timed region
    stage_sum = {}
    nst = min(args.steps, 20)
    for _ in range(nst):
        el.process_device(L, R, D1, D2)
        for k, v in el.stage_times().items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    torch.cuda.synchronize()
    stages = {k: v / nst for k, v in stage_sum.items()}
    value = args.steps * mpxd / wall
    host = None
    if not args.no_host:
        # the reference's own call: pageable host buffers in and out
        h1 = np.empty((H, W), np.float32)
        h2 = np.empty((H, W), np.float32)
        for _ in range(3):
            el.process(left, right, h1, h2)
        nh = max(10, args.steps // 2)
        th = []
        for _ in range(nh):
            tc = time.perf_counter()
            el.process(left, right, h1, h2)
            th.append(time.perf_counter() - tc)
        th = np.array(th) * 1e3
        host = {"pairs_per_s": round(1e3 / th.mean(), 2), "mpx_disp_per_s": round(mpxd * 1e3 / th.mean(), 1),
                "p50_ms": round(float(np.median(th)), 4), "p90_ms": round(float(np.percentile(th, 90)), 4),
                "calls": nh, "buffers": "pageable host (elas_hip_process, PCIe copies included)",
                "same_maps": bool(np.array_equal(h1, D1.cpu().numpy()) and np.array_equal(h2, D2.cpu().numpy()))}

    wc = work_counts(oracle, left, right, dmax)
    kernels = kernel_rooflines(el, wc, W, H, dmax, warm_ms=args.kernel_warm_ms)
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    d = kernels[dom]
    roofline = {"bound": d["bound"], "achieved": round(d["achieved"], 4), "peak": round(d["peak"], 3),
                "unit": d["unit"], "frac": round(d["frac"], 5), "traffic": traffic_of(dom, args.workload, W, H, dmax),
                "kernel": dom, "ms_per_launch": round(d["ms"], 5)}

    res = {
        "metric": "Mpx*disp/s", "value": round(value, 2), "unit": "Mpx*disp/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "elas_plain_%dx%d_d%d_device_resident" % (W, H, dmax + 1), "width": W, "height": H,
                   "disp_max": dmax, "pairs_per_step": 1, "host_delaunay": "included"},
        "pairs_per_s": round(args.steps / wall, 2),
        "per_call_ms": {"p50": round(float(np.median(per_call)), 4), "p90": round(float(np.percentile(per_call, 90)), 4),
                        "min": round(float(per_call.min()), 4), "max": round(float(per_call.max()), 4)},
        "host_delaunay": {"split_depth": pkg.tri_split_depth(int(el.shape_info()["num_support"])),
                          "support_points": int(el.shape_info()["num_support"])},
        "host_buffers": host,
        "stages_ms": {k: round(v, 5) for k, v in stages.items()},
        "gpu_phase_ms_per_pair": round(stages["phase_a"] + stages["phase_b"], 5),
        "launch_mode": "hip graphs per shape bucket (stages inside a graph read -1)",
        "dense_tiles": dict(zip(("cells_per_wave", "steps_per_chunk", "tuned_us"),
                                el.dense_tiles(W, H, dmax))) | {"forced": bool(args.dense_tiles)},
        "roofline": roofline,
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
        "work": wc,
    }
    pool = None
    if not args.no_batch or not args.no_serve:
        pool = synth.batch_pairs(64)
    if not args.no_batch:
        res["batch_vs_plain"] = batch_vs_plain(pkg, synth, pool=pool)
    if not args.no_serve:
        if args.serve_quick:
            res["serving"] = serving(pkg, synth, rates=(100, 400, 800), n_req=400, pool=pool, knee=())
        else:
            res["serving"] = serving(pkg, synth, pool=pool)
    if not args.no_cpu:
        n, el_s, cores, el_all = cpu_baseline(oracle, left, right, dmax, args.cpu_seconds)
        model = ""
        try:
            model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
        except Exception:
            pass
        res["host_cpu"] = {"model": model, "logical_cpus": os.cpu_count()}
        res["cpu_baseline"] = {"value": round(cores * mpxd / el_all, 3), "unit": "Mpx*disp/s", "cores": cores,
                               "kind": "port", "sample": "%d full pipelines of the same %dx%d d%d pair at once, one "
                               "per logical CPU (oracle, %d processes, %.1f s)" % (cores, W, H, dmax + 1, cores, el_all),
                               "single_core": {"value": round(n * mpxd / el_s, 3), "cores": 1,
                                               "sample": "%d full pipelines back to back (%.1f s)" % (n, el_s)}}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
