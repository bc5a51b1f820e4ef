"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle restates the reference ELAS kernels on the CPU (elas_oracle.cpp)
and checks the HIP path.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package; the product never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libelas_oracle.so")


class Params(ctypes.Structure):
    """Field order and types of Elas::parameters (src/Elas.h:15-37)."""
    _fields_ = [
        ("disp_min", ctypes.c_int32), ("disp_max", ctypes.c_int32),
        ("support_threshold", ctypes.c_float), ("support_texture", ctypes.c_int32),
        ("candidate_stepsize", ctypes.c_int32), ("incon_window_size", ctypes.c_int32),
        ("incon_threshold", ctypes.c_int32), ("incon_min_support", ctypes.c_int32),
        ("add_corners", ctypes.c_bool), ("grid_size", ctypes.c_int32),
        ("beta", ctypes.c_float), ("gamma", ctypes.c_float), ("sigma", ctypes.c_float),
        ("sradius", ctypes.c_float), ("match_texture", ctypes.c_int32),
        ("lr_threshold", ctypes.c_int32), ("speckle_sim_threshold", ctypes.c_float),
        ("speckle_size", ctypes.c_int32), ("ipol_gap_width", ctypes.c_int32),
        ("filter_median", ctypes.c_bool), ("filter_adaptive_mean", ctypes.c_bool),
        ("postprocess_only_left", ctypes.c_bool), ("subsampling", ctypes.c_bool),
    ]


# Effective values of the reference kernels (src/Kernels/KernelsCommon.h:8-27),
# post-processing off (src/Elas.cu:107-119 ends at the LR check).
EFFECTIVE = dict(disp_min=0, disp_max=255, support_threshold=0.85, support_texture=10,
                 candidate_stepsize=5, incon_window_size=5, incon_threshold=5, incon_min_support=5,
                 add_corners=False, grid_size=20, beta=0.02, gamma=3.0, sigma=1.0, sradius=2.0,
                 match_texture=0, lr_threshold=2, speckle_sim_threshold=1.0, speckle_size=0,
                 ipol_gap_width=0, filter_median=False, filter_adaptive_mean=False,
                 postprocess_only_left=False, subsampling=False)


def params(**kw):
    d = dict(EFFECTIVE)
    d.update(kw)
    return Params(**d)


class _Dump(ctypes.Structure):
    P = ctypes.c_void_p
    _fields_ = [("desc0", P), ("desc1", P), ("can_left", P), ("can_lr", P), ("can_incon", P),
                ("can_final", P), ("sp0", P), ("sp1", P), ("grid0", P), ("grid1", P),
                ("tri0", P), ("tri1", P), ("tri_cap", ctypes.c_int32), ("mean0", P), ("mean1", P),
                ("d1_raw", P), ("d2_raw", P), ("nsp", ctypes.c_int32), ("ntri0", ctypes.c_int32),
                ("ntri1", ctypes.c_int32), ("work_support", ctypes.c_int64),
                ("work_dense", ctypes.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError("oracle library not built: run oracle/build_oracle.py (%s)" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH)
        for name in ("eo_compact", "eo_triangulate", "eo_process"):
            getattr(_lib, name).restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _img(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    assert a.ndim == 2
    return a


def prior(prm, delta):
    return lib().eo_prior(ctypes.byref(prm), int(delta))


def plane_radius(prm):
    return lib().eo_plane_radius(ctypes.byref(prm))


def sobel(img):
    img = _img(img)
    H, W = img.shape
    du = np.zeros((H, W), np.uint8)
    dv = np.zeros((H, W), np.uint8)
    lib().eo_sobel(_p(img), W, H, _p(du), _p(dv))
    return du, dv


def descriptor(img):
    img = _img(img)
    H, W = img.shape
    desc = np.zeros((H, W, 16), np.uint8)
    lib().eo_descriptor(_p(img), W, H, _p(desc))
    return desc


def support_pass(desc0, desc1, prm, right, can):
    """One support pass in place on can (int32 [ch, cw]); returns work count."""
    H, W = desc0.shape[:2]
    work = ctypes.c_int64(0)
    lib().eo_support_pass(_p(desc0), _p(desc1), W, H, ctypes.byref(prm), int(right), _p(can),
                          ctypes.byref(work))
    return work.value


def incon_filter(can, prm):
    ch, cw = can.shape
    out = np.empty_like(can)
    lib().eo_incon_filter(_p(np.ascontiguousarray(can, np.int32)), _p(out), cw, ch, ctypes.byref(prm))
    return out


def redun_filter(can):
    can = np.ascontiguousarray(can, np.int32).copy()
    ch, cw = can.shape
    lib().eo_redun_filter(_p(can), cw, ch)
    return can


def compact(can, step):
    can = np.ascontiguousarray(can, np.int32)
    ch, cw = can.shape
    sp0 = np.zeros((cw * ch + 1, 2), np.float32)
    sp1 = np.zeros((cw * ch + 1, 2), np.float32)
    n = lib().eo_compact(_p(can), cw, ch, step, _p(sp0), _p(sp1))
    return sp0[:n].copy(), sp1[:n].copy()


def grid_dims(W, H, prm):
    gw, gh = ctypes.c_int(), ctypes.c_int()
    lib().eo_grid_dims(W, H, ctypes.byref(prm), ctypes.byref(gw), ctypes.byref(gh))
    return gw.value, gh.value


def grid(can, W, H, prm):
    can = np.ascontiguousarray(can, np.int32)
    ch, cw = can.shape
    gw, gh = grid_dims(W, H, prm)
    g0 = np.zeros((gh, gw, prm.disp_max + 2), np.int32)
    g1 = np.zeros_like(g0)
    lib().eo_grid(_p(can), cw, ch, W, H, ctypes.byref(prm), _p(g0), _p(g1))
    return g0, g1


def triangulate(sp):
    sp = np.ascontiguousarray(sp, np.float32)
    cap = 2 * len(sp) + 8
    tri = np.zeros((cap, 3), np.int32)
    n = lib().eo_triangulate(_p(sp), len(sp), _p(tri), cap)
    if n < 0:
        raise ValueError("triangulation needs at least 3 points")
    return tri[:n].copy()


def add_corners(sp0, sp1, W, H):
    """The support points with the corner points of add_corners appended
    (eo_add_corners): (sp0, sp1) as float32 [n][2] arrays."""
    n = len(sp0)
    a = np.zeros((n + 6, 2), np.float32)
    b = np.zeros((n + 6, 2), np.float32)
    a[:n], b[:n] = sp0, sp1
    m = lib().eo_add_corners(_p(a), _p(b), n, int(W), int(H))
    return a[:m].copy(), b[:m].copy()


def plane(can, step, sp, tri, right):
    can = np.ascontiguousarray(can, np.int32)
    abc = np.zeros(3, np.float32)
    t = np.ascontiguousarray(tri, np.int32)
    lib().eo_plane(_p(can), can.shape[1], step, _p(np.ascontiguousarray(sp, np.float32)), _p(t),
                   int(right), _p(abc))
    return abc


def raster(can, sp0, tri, right, W, H, prm):
    can = np.ascontiguousarray(can, np.int32)
    mean = np.full((H, W), -1.0, np.float32)
    tri = np.ascontiguousarray(tri, np.int32)
    lib().eo_raster(_p(can), can.shape[1], _p(np.ascontiguousarray(sp0, np.float32)), _p(tri),
                    len(tri), int(right), W, H, ctypes.byref(prm), _p(mean))
    return mean


def dense(grid_, desc0, desc1, mean, right, prm):
    H, W = mean.shape
    disp = np.empty((H, W), np.float32)
    work = ctypes.c_int64(0)
    lib().eo_dense(_p(np.ascontiguousarray(grid_, np.int32)), _p(desc0), _p(desc1),
                   _p(np.ascontiguousarray(mean, np.float32)), int(right), W, H, ctypes.byref(prm),
                   _p(disp), ctypes.byref(work))
    return disp, work.value


def lr_check(D1, D2, prm):
    D1 = np.ascontiguousarray(D1, np.float32).copy()
    D2 = np.ascontiguousarray(D2, np.float32).copy()
    H, W = D1.shape
    lib().eo_lr(_p(D1), _p(D2), W, H, ctypes.byref(prm))
    return D1, D2


def postprocess(D, prm):
    D = np.ascontiguousarray(D, np.float32).copy()
    H, W = D.shape
    lib().eo_postprocess(_p(D), W, H, ctypes.byref(prm))
    return D


def speckle(D, sim, size):
    D = np.ascontiguousarray(D, np.float32).copy()
    lib().eo_speckle(_p(D), D.shape[1], D.shape[0], ctypes.c_float(sim), int(size))
    return D


def gap_interpolation(D, gap):
    D = np.ascontiguousarray(D, np.float32).copy()
    lib().eo_gap_interpolation(_p(D), D.shape[1], D.shape[0], int(gap))
    return D


def adaptive_mean(D):
    D = np.ascontiguousarray(D, np.float32).copy()
    lib().eo_adaptive_mean(_p(D), D.shape[1], D.shape[0])
    return D


def median(D):
    D = np.ascontiguousarray(D, np.float32).copy()
    lib().eo_median(_p(D), D.shape[1], D.shape[0])
    return D


def process(left, right, prm=None, dump=True):
    """Whole pipeline.  Returns dict with D1, D2, status and (dump=True) every
    intermediate: desc0/1, can_left/lr/incon/final, sp0/sp1, grid0/1,
    tri0/tri1, mean0/1, d1_raw/d2_raw, work_support, work_dense."""
    left, right = _img(left), _img(right)
    H, W = left.shape
    prm = prm or params()
    D1 = np.empty((H, W), np.float32)
    D2 = np.empty((H, W), np.float32)
    out = {"D1": D1, "D2": D2}
    dp = None
    if dump:
        step = prm.candidate_stepsize
        cw, ch = W // step, H // step
        gw, gh = grid_dims(W, H, prm)
        cap = 2 * (cw * ch + 8) + 8
        bufs = dict(desc0=np.zeros((H, W, 16), np.uint8), desc1=np.zeros((H, W, 16), np.uint8),
                    can_left=np.zeros((ch, cw), np.int32), can_lr=np.zeros((ch, cw), np.int32),
                    can_incon=np.zeros((ch, cw), np.int32), can_final=np.zeros((ch, cw), np.int32),
                    sp0=np.zeros((cw * ch + 8, 2), np.float32), sp1=np.zeros((cw * ch + 8, 2), np.float32),
                    grid0=np.zeros((gh, gw, prm.disp_max + 2), np.int32),
                    grid1=np.zeros((gh, gw, prm.disp_max + 2), np.int32),
                    tri0=np.zeros((cap, 3), np.int32), tri1=np.zeros((cap, 3), np.int32),
                    mean0=np.full((H, W), -1, np.float32), mean1=np.full((H, W), -1, np.float32),
                    d1_raw=np.full((H, W), -1, np.float32), d2_raw=np.full((H, W), -1, np.float32))
        dp = _Dump(**{k: _p(v) for k, v in bufs.items()}, tri_cap=cap)
    st = lib().eo_process(_p(left), _p(right), W, H, ctypes.byref(prm), _p(D1), _p(D2),
                          ctypes.byref(dp) if dp is not None else None)
    out["status"] = st
    if dump:
        out.update(bufs)
        out["nsp"], out["ntri0"], out["ntri1"] = dp.nsp, dp.ntri0, dp.ntri1
        out["sp0"] = bufs["sp0"][:dp.nsp]
        out["sp1"] = bufs["sp1"][:dp.nsp]
        out["tri0"] = bufs["tri0"][:dp.ntri0]
        out["tri1"] = bufs["tri1"][:dp.ntri1]
        out["work_support"], out["work_dense"] = dp.work_support, dp.work_dense
    return out


def _maps_job(args):
    left, right, dmax = args
    out = process(left, right, params(disp_max=dmax), dump=False)
    return out["D1"], out["D2"], out["status"]


def process_many(pairs, workers=16):
    """[(D1, D2, status)] of the whole pipeline for each (left, right,
    disp_max), computed in up to `workers` fresh processes (Triangle keeps its
    state in globals, so one process runs one pipeline at a time).  The
    processes are spawned, not forked, so a caller that has initialised a GPU
    hands none of its state to them."""
    import concurrent.futures as cf
    import multiprocessing as mp
    workers = max(1, min(workers, len(pairs), os.cpu_count() or 1))
    if workers == 1:
        return [_maps_job(p) for p in pairs]
    with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_maps_job, pairs))
