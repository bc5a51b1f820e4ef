"""MI355X-native ELAS stereo path: Python front end over libelas_hip.so.

Mirrors the reference's interface (class Elas, src/Elas.h:11-167):
``Elas(params).process(left, right) -> (D1, D2)`` with the same parameter
struct, plus the batched and device-resident forms.  Every call goes to the
gfx950 kernels through the C ABI of include/elas_hip.h; there is no CPU
fallback: a missing library or device raises.

The directory name is not a Python identifier; load it with
``importlib`` (see ``__graft_entry__.load_package``) under the name
``cuda_stereo_matching_amd``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libelas_hip.so")
CLI_PATH = os.path.join(HERE, "lib", "elas_cli")  # command-line tool (src/Main.cu counterpart)

STATUS = {0: "ok", 1: "invalid argument", 2: "HIP error", 3: "out of memory",
          4: "too few support points", 5: "no gfx950 device", 6: "unsupported", 7: "busy", 8: "I/O error"}
ELAS_OK, ELAS_ERR_TOO_FEW_POINTS, ELAS_ERR_BUSY, ELAS_ERR_IO = 0, 4, 7, 8

STAGES = ["desc0", "desc1", "can_left", "can_lr", "can_incon", "can_final", "sp0", "sp1", "grid0",
          "grid1", "tri0", "tri1", "mean0", "mean1", "d1_raw", "d2_raw", "D1", "D2"]
KERNELS = ["descriptor", "support", "dense", "lr"]  # elas_kernel order
TIMERS = ["descriptor", "support", "filter", "compact", "grid", "triangle", "plane", "dense", "lr",
          "post", "total", "phase_a", "phase_b", "sp_download", "tri_wall", "tri_job",
          "host_a", "host_b", "gpu_span", "ab_gap"]


class ElasParams(ctypes.Structure):
    """elas_params: same field order and types as Elas::parameters (src/Elas.h:15-37)."""
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

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ShapeInfo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("width", "height", "can_width", "can_height", "grid_width",
                                               "grid_height", "num_support", "num_tri0", "num_tri1",
                                               "status")]


# Exported symbols and their signatures (include/elas_hip.h).
_P = ctypes.c_void_p
class ServeConfig(ctypes.Structure):
    """elas_serve_config (include/elas_hip.h)."""
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_chunk_pairs", ctypes.c_int32),
                ("slots", ctypes.c_int32)]


class ServeStats(ctypes.Structure):
    """elas_serve_stats (include/elas_hip.h)."""
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("chunks", ctypes.c_int64),
                ("max_chunk_pairs", ctypes.c_int64), ("max_in_flight_pairs", ctypes.c_int64),
                ("slot_capacity_bytes", ctypes.c_uint64), ("peak_in_flight_bytes", ctypes.c_uint64)]


SIGNATURES = {
    "elas_params_default": (None, [ctypes.POINTER(ElasParams)]),
    "elas_params_declared": (None, [ctypes.POINTER(ElasParams)]),
    "elas_hip_create": (ctypes.c_int, [ctypes.POINTER(ElasParams), ctypes.c_int, ctypes.POINTER(_P)]),
    "elas_hip_destroy": (ctypes.c_int, [_P]),
    "elas_hip_set_params": (ctypes.c_int, [_P, ctypes.POINTER(ElasParams)]),
    "elas_hip_last_error": (ctypes.c_char_p, [_P]),
    "elas_hip_stream": (_P, [_P]),
    "elas_hip_process": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "elas_hip_process_device": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32]),
    "elas_hip_process_batch": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _P, _P, _P, _P]),
    "elas_hip_process_batch_device": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _P, _P, _P, _P]),
    "elas_hip_set_pipeline": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "elas_hip_set_graphs": (ctypes.c_int, [_P, ctypes.c_int32]),
    "elas_hip_time_kernel": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    "elas_hip_serve_start": (ctypes.c_int, [_P, _P]),
    "elas_hip_submit": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "elas_hip_poll": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_float)]),
    "elas_hip_serve_stop": (ctypes.c_int, [_P]),
    "elas_hip_serve_stats": (ctypes.c_int, [_P, _P]),
    "elas_hip_set_debug": (ctypes.c_int, [_P, ctypes.c_int]),
    "elas_hip_debug_fetch": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_size_t)]),
    "elas_hip_shape_info": (ctypes.c_int, [_P, ctypes.POINTER(ShapeInfo)]),
    "elas_hip_stage_times": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_int32]),
    "elas_hip_triangulate": (ctypes.c_int, [_P, ctypes.c_int32, _P, ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_int64)]),
    "elas_hip_set_autotune": (ctypes.c_int, [_P, ctypes.c_int32]),
    "elas_hip_process_device_noalloc": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32]),
    "elas_hip_set_tri_split": (ctypes.c_int, [ctypes.c_int32]),
    "elas_hip_tri_split_depth": (ctypes.c_int, [ctypes.c_int32]),
    "elas_io_read_pgm": (ctypes.c_int, [ctypes.c_char_p, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_int32)]),
    "elas_io_write_pgm": (ctypes.c_int, [ctypes.c_char_p, _P, ctypes.c_int32, ctypes.c_int32]),
    "elas_io_write_pfm": (ctypes.c_int, [ctypes.c_char_p, _P, ctypes.c_int32, ctypes.c_int32]),
    "elas_io_disparity_to_u8": (ctypes.c_int, [_P, _P, ctypes.c_int64]),
    "elas_hip_set_dense_tiles": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "elas_hip_set_desc_rows": (ctypes.c_int, [_P, ctypes.c_int32]),
    "elas_hip_desc_rows": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float)]),
    "elas_hip_dense_tiles": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                            ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load libelas_hip.so (raises if it has not been built).

    torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    Importing torch first makes the library bind to that runtime, so a
    process holds one HIP runtime and torch tensors can be passed to
    process_device; loading the library first would leave torch a second
    runtime that cannot open the device."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(path):
        raise RuntimeError("libelas_hip.so is not built (%s): run __graft_entry__.build()" % path)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _torch_sync(t):
    """The library enqueues on its own streams: work torch queued on the
    tensors' current stream has to finish first (a query first: an idle
    stream needs no synchronisation call)."""
    import torch
    s = torch.cuda.current_stream(t.device)
    if not s.query():
        s.synchronize()


def default_params(**overrides):
    """Effective reference defaults (src/Kernels/KernelsCommon.h:8-27) with overrides."""
    p = ElasParams()
    load_library().elas_params_default(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def declared_params(**overrides):
    """The initialisers of src/Elas.h:42-66 with overrides."""
    p = ElasParams()
    load_library().elas_params_declared(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


class ElasError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "error"), code, msg))
        self.code = code


def triangulate(points):
    """Delaunay triangles of [n][2] float points exactly as the pipeline's host
    step computes them (the reference's Triangle, switches "zQB",
    src/Elas.cu:278-328): an [ntri][3] int32 array of zero-based indices.
    Host only; no device needed."""
    lib = load_library()
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 2)
    cap = 2 * len(pts) + 8
    tri = np.zeros((cap, 3), np.int32)
    n = ctypes.c_int64()
    st = lib.elas_hip_triangulate(pts.ctypes.data, len(pts), tri.ctypes.data, cap, ctypes.byref(n))
    if st != ELAS_OK:
        raise ElasError(st, "triangulate: status %d" % st)
    return tri[:min(n.value, cap)].copy()


def set_tri_split(max_depth=-1):
    """Process-wide split of each mesh's divide-and-conquer over the host
    Triangle threads: up to 2**max_depth concurrent subtrees (0: serial, -1:
    automatic).  Meshes are the same under every setting."""
    st = load_library().elas_hip_set_tri_split(int(max_depth))
    if st != ELAS_OK:
        raise ElasError(st, "set_tri_split(%r)" % (max_depth,))


def tri_split_depth(n):
    """Split depth a mesh of n points gets under the current setting."""
    return int(load_library().elas_hip_tri_split_depth(int(n)))


def read_pgm(path):
    """uint8 [H, W] image of a binary PGM (the reference's loadPGM,
    src/Image.h:129-160); raises ElasError(ELAS_ERR_IO) on a malformed file."""
    lib = load_library()
    w, h = ctypes.c_int32(), ctypes.c_int32()
    p = os.fsencode(path)
    st = lib.elas_io_read_pgm(p, None, 0, ctypes.byref(w), ctypes.byref(h))
    if st != ELAS_OK:
        raise ElasError(st, "read_pgm: %s" % path)
    img = np.empty((h.value, w.value), np.uint8)
    st = lib.elas_io_read_pgm(p, img.ctypes.data, img.size, ctypes.byref(w), ctypes.byref(h))
    if st != ELAS_OK:
        raise ElasError(st, "read_pgm: %s" % path)
    return img


def write_pgm(path, img):
    """Binary PGM of a uint8 [H, W] image (savePGM, src/Image.h:162-169)."""
    img = _u8(img)
    st = load_library().elas_io_write_pgm(os.fsencode(path), img.ctypes.data, img.shape[1], img.shape[0])
    if st != ELAS_OK:
        raise ElasError(st, "write_pgm: %s" % path)


def write_pfm(path, disp):
    """Float map as PFM (scale -1: little endian, bottom row first)."""
    d = np.ascontiguousarray(disp, dtype=np.float32)
    st = load_library().elas_io_write_pfm(os.fsencode(path), d.ctypes.data, d.shape[1], d.shape[0])
    if st != ELAS_OK:
        raise ElasError(st, "write_pfm: %s" % path)


def disparity_to_u8(disp):
    """The 8-bit view of a disparity map that the reference displays
    (src/Main.cu:14-27): (uint8)max(255 * d / max(d), 0)."""
    d = np.ascontiguousarray(disp, dtype=np.float32)
    out = np.empty(d.shape, np.uint8)
    st = load_library().elas_io_disparity_to_u8(d.ctypes.data, out.ctypes.data, d.size)
    if st != ELAS_OK:
        raise ElasError(st, "disparity_to_u8")
    return out


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    if a.ndim != 2:
        raise ValueError("images must be 2-D uint8 arrays")
    return a


class Elas:
    """Drop-in for the reference's class Elas on one gfx950 GPU."""

    def __init__(self, params=None, device=0):
        self._lib = load_library()
        self.params = params if params is not None else default_params()
        self._ctx = ctypes.c_void_p()
        st = self._lib.elas_hip_create(ctypes.byref(self.params), int(device), ctypes.byref(self._ctx))
        if st != 0:
            raise ElasError(st, "elas_hip_create(device=%d)" % device)
        self.status = 0

    def close(self):
        if self._ctx:
            self._lib.elas_hip_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, allow_few=False):
        self.status = st
        if st == 0 or (allow_few and st == ELAS_ERR_TOO_FEW_POINTS):
            return st
        raise ElasError(st, (self._lib.elas_hip_last_error(self._ctx) or b"").decode())

    def last_error(self):
        return (self._lib.elas_hip_last_error(self._ctx) or b"").decode()

    def set_params(self, params):
        self._check(self._lib.elas_hip_set_params(self._ctx, ctypes.byref(params)))
        self.params = params

    def process(self, left, right, D1=None, D2=None):
        """Elas::process (src/Elas.h:76-83): host uint8 [H,W] in, float32 D1/D2 out.
        Fewer than 3 support points returns all -1 maps with status 4 (the
        reference exits)."""
        left, right = _u8(left), _u8(right)
        if left.shape != right.shape:
            raise ValueError("images must have the same size")
        H, W = left.shape
        D1 = np.empty((H, W), np.float32) if D1 is None else D1
        D2 = np.empty((H, W), np.float32) if D2 is None else D2
        dims = (ctypes.c_int32 * 3)(W, H, W)
        st = self._lib.elas_hip_process(self._ctx, left.ctypes.data, right.ctypes.data, D1.ctypes.data,
                                        D2.ctypes.data, ctypes.cast(dims, ctypes.c_void_p))
        self._check(st, allow_few=True)
        return D1, D2

    def process_device(self, left, right, D1, D2):
        """Device-resident form: torch uint8 [H,W] cuda tensors in, float32 out."""
        import torch
        H, W = int(left.shape[0]), int(left.shape[1])
        for t, dt in ((left, torch.uint8), (right, torch.uint8), (D1, torch.float32), (D2, torch.float32)):
            if t.dtype is not dt or not t.is_cuda or t.dim() != 2 or t.shape[0] != H or t.shape[1] != W \
                    or not t.is_contiguous():
                raise ValueError("process_device expects contiguous cuda tensors of shape [H, W]")
        _torch_sync(left)
        st = self._lib.elas_hip_process_device(self._ctx, left.data_ptr(), right.data_ptr(), D1.data_ptr(),
                                               D2.data_ptr(), W, H)
        return self._check(st, allow_few=True)

    def process_device_noalloc(self, left, right, D1, D2):
        """process_device without any device or pinned allocation: the
        intermediates live in the library's per-GPU scratch (128 MiB), so the
        pair must fit it (ELAS_ERR_UNSUPPORTED otherwise); calls that overlap
        on one GPU race."""
        import torch
        H, W = int(left.shape[0]), int(left.shape[1])
        for t, dt in ((left, torch.uint8), (right, torch.uint8), (D1, torch.float32), (D2, torch.float32)):
            if t.dtype is not dt or not t.is_cuda or t.dim() != 2 or t.shape[0] != H or t.shape[1] != W \
                    or not t.is_contiguous():
                raise ValueError("process_device_noalloc expects contiguous cuda tensors of shape [H, W]")
        _torch_sync(left)
        st = self._lib.elas_hip_process_device_noalloc(self._ctx, left.data_ptr(), right.data_ptr(), D1.data_ptr(),
                                                       D2.data_ptr(), W, H)
        return self._check(st, allow_few=True)

    def time_kernel(self, kernel, reps=20):
        """Average ms per launch of one kernel of the last call, re-launched
        `reps` times back to back (kernel: one of KERNELS)."""
        ms = ctypes.c_float(0)
        self._check(self._lib.elas_hip_time_kernel(self._ctx, KERNELS.index(kernel), int(reps), ctypes.byref(ms)))
        return float(ms.value)

    # ---- serving -----------------------------------------------------------
    def serve_start(self, hbm_budget_bytes=0, max_chunk_pairs=0, slots=0):
        """Start the context's scheduler thread (continuous batching over the
        pipeline slots, admission by HBM budget).  0 = library default."""
        cfg = ServeConfig(int(hbm_budget_bytes), int(max_chunk_pairs), int(slots))
        self._check(self._lib.elas_hip_serve_start(self._ctx, ctypes.byref(cfg)))
        self._inflight = {}

    def submit(self, left, right, D1, D2, disp_max=None):
        """Queue one pair; returns a ticket.  numpy (host) or torch cuda
        (device-resident) buffers; they are kept referenced until polled done.
        Host buffers should be pinned for the copies to overlap."""
        dev = hasattr(left, "is_cuda") and left.is_cuda
        if dev:
            H, W = int(left.shape[0]), int(left.shape[1])
            ptr = lambda t: t.data_ptr()
            _torch_sync(left)
        else:
            H, W = left.shape
            for a, dt in ((left, np.uint8), (right, np.uint8), (D1, np.float32), (D2, np.float32)):
                if a.dtype != dt or not a.flags.c_contiguous or a.shape != (H, W):
                    raise ValueError("submit expects contiguous [H, W] uint8 images and float32 maps")
            ptr = lambda a: a.ctypes.data
        t = ctypes.c_int64(0)
        dm = -1 if disp_max is None else int(disp_max)
        self._check(self._lib.elas_hip_submit(self._ctx, ptr(left), ptr(right), ptr(D1), ptr(D2), W, H, dm,
                                              1 if dev else 0, ctypes.byref(t)))
        self._inflight[t.value] = (left, right, D1, D2)
        return t.value

    def poll(self, ticket):
        """None while queued or in flight, else (status, latency_ms)."""
        st = ctypes.c_int32(0)
        lat = ctypes.c_float(0)
        r = self._lib.elas_hip_poll(self._ctx, int(ticket), ctypes.byref(st), ctypes.byref(lat))
        if r == ELAS_ERR_BUSY:
            return None
        self._check(r)
        self._inflight.pop(int(ticket), None)
        return int(st.value), float(lat.value)

    def serve_stats(self):
        s = ServeStats()
        self._check(self._lib.elas_hip_serve_stats(self._ctx, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def serve_stop(self):
        """Drain the queue and stop the scheduler thread."""
        self._check(self._lib.elas_hip_serve_stop(self._ctx))

    def set_graphs(self, on=True):
        """Replay phases A and B as HIP graphs per shape bucket (default), or
        launch kernel by kernel with per-stage timing events."""
        self._check(self._lib.elas_hip_set_graphs(self._ctx, 1 if on else 0))

    def set_autotune(self, on=True):
        """Time every dense tiling on the first pair of each new (W, H,
        disp_max) and keep the fastest (default), or use the heuristic."""
        self._check(self._lib.elas_hip_set_autotune(self._ctx, 1 if on else 0))

    def set_dense_tiles(self, cells_per_wave=0, steps_per_chunk=0):
        """Force the dense tiling (cells per wave 1 or 2, steps per chunk;
        0 = the heuristic's); (0, 0) restores automatic tiling."""
        self._check(self._lib.elas_hip_set_dense_tiles(self._ctx, int(cells_per_wave), int(steps_per_chunk)))

    def dense_tiles(self, W, H, disp_max):
        """(cells per wave, steps per chunk, tuned microseconds or -1) a
        shape would be matched with now."""
        c, s, us = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
        self._check(self._lib.elas_hip_dense_tiles(self._ctx, int(W), int(H), int(disp_max), ctypes.byref(c),
                                                   ctypes.byref(s), ctypes.byref(us)))
        return c.value, s.value, us.value

    def set_desc_rows(self, rows=0):
        """Force the descriptor strip height (6, 12 or 18 rows; 0 = automatic)."""
        self._check(self._lib.elas_hip_set_desc_rows(self._ctx, int(rows)))

    def desc_rows(self, W, H, disp_max):
        """(strip rows, tuned microseconds or -1) a shape's descriptor would use now."""
        r, us = ctypes.c_int32(), ctypes.c_float()
        self._check(self._lib.elas_hip_desc_rows(self._ctx, int(W), int(H), int(disp_max), ctypes.byref(r),
                                                 ctypes.byref(us)))
        return r.value, us.value

    def set_pipeline(self, slots=3, chunk=0):
        """Batches are cut into chunks of `chunk` pairs (0: one chunk) that
        move through `slots` pipeline slots, overlapping the GPU stages of one
        chunk with the host Delaunay step of another."""
        self._check(self._lib.elas_hip_set_pipeline(self._ctx, int(slots), int(chunk)))

    def process_batch_device(self, lefts, rights, D1s, D2s, disp_max=None):
        """Device-resident batch: lists of torch cuda tensors (uint8 [H,W] in,
        float32 [H,W] out), shapes may differ per pair; disp_max is None (the
        context's) or a per-pair list.  Returns the per-pair statuses."""
        n = len(lefts)
        if not (len(rights) == len(D1s) == len(D2s) == n) or n == 0:
            raise ValueError("need n > 0 matching lists")
        dims = (ctypes.c_int32 * (3 * n))()
        for i in range(n):
            H, W = int(lefts[i].shape[0]), int(lefts[i].shape[1])
            for t, dt in ((lefts[i], "uint8"), (rights[i], "uint8"), (D1s[i], "float32"), (D2s[i], "float32")):
                if not t.is_cuda or not t.is_contiguous() or str(t.dtype) != "torch." + dt or tuple(t.shape) != (H, W):
                    raise ValueError("pair %d: expected contiguous cuda tensors of shape [H, W]" % i)
            dims[3 * i], dims[3 * i + 1] = W, H
            dims[3 * i + 2] = -1 if disp_max is None or disp_max[i] is None else int(disp_max[i])
        arr = lambda xs: (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
        stat = (ctypes.c_int32 * n)()
        _torch_sync(lefts[0])
        st = self._lib.elas_hip_process_batch_device(self._ctx, n, ctypes.cast(arr(lefts), _P),
                                                     ctypes.cast(arr(rights), _P), ctypes.cast(arr(D1s), _P),
                                                     ctypes.cast(arr(D2s), _P), ctypes.cast(dims, _P),
                                                     ctypes.cast(stat, _P))
        self._check(st)
        return [int(stat[i]) for i in range(n)]

    def process_batch(self, pairs):
        """pairs: list of (left, right, disp_max or None).  One launch per stage
        over all pairs.  Returns list of (D1, D2, status)."""
        n = len(pairs)
        ls = [_u8(p[0]) for p in pairs]
        rs = [_u8(p[1]) for p in pairs]
        d1 = [np.empty(l.shape, np.float32) for l in ls]
        d2 = [np.empty(l.shape, np.float32) for l in ls]
        arr = lambda xs: (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
        dims = (ctypes.c_int32 * (3 * n))()
        for i, (l, p) in enumerate(zip(ls, pairs)):
            dims[3 * i], dims[3 * i + 1] = l.shape[1], l.shape[0]
            dims[3 * i + 2] = -1 if len(p) < 3 or p[2] is None else int(p[2])
        stat = (ctypes.c_int32 * n)()
        st = self._lib.elas_hip_process_batch(self._ctx, n, ctypes.cast(arr(ls), _P), ctypes.cast(arr(rs), _P),
                                              ctypes.cast(arr(d1), _P), ctypes.cast(arr(d2), _P),
                                              ctypes.cast(dims, _P), ctypes.cast(stat, _P))
        self._check(st)
        return [(d1[i], d2[i], int(stat[i])) for i in range(n)]

    def set_debug(self, on=True, tiny_lists=False, gate=True, short_gate=False):
        """Keep every intermediate of the next calls for fetch(); tiny_lists
        caps the raster's per-tile triangle lists at one entry (a test of the
        overflow path, where a tile tests every triangle of the mesh); gate
        False launches phase B after the host Delaunay step instead of
        queueing it behind a gate during that step; short_gate bounds the
        gate's wait at one tick (a test of its failure path)."""
        self._check(self._lib.elas_hip_set_debug(
            self._ctx, (1 if on else 0) | (2 if tiny_lists else 0) | (0 if gate else 4) | (8 if short_gate else 0)))

    def shape_info(self):
        info = ShapeInfo()
        self._check(self._lib.elas_hip_shape_info(self._ctx, ctypes.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def stage_times(self):
        ms = (ctypes.c_float * len(TIMERS))()
        self._check(self._lib.elas_hip_stage_times(self._ctx, ms, len(TIMERS)))
        return dict(zip(TIMERS, list(ms)))

    def fetch(self, stage):
        """Intermediate of the last call (debug mode), shaped like the reference buffer."""
        k = STAGES.index(stage)
        sz = ctypes.c_size_t(0)
        self._check(self._lib.elas_hip_debug_fetch(self._ctx, k, None, 0, ctypes.byref(sz)))
        info = self.shape_info()
        W, H = info["width"], info["height"]
        cw, ch = info["can_width"], info["can_height"]
        gw, gh = info["grid_width"], info["grid_height"]
        if stage.startswith("desc"):
            out = np.zeros((H, W, 16), np.uint8)
        elif stage.startswith("can"):
            out = np.zeros((ch, cw), np.int32)
        elif stage.startswith("sp"):
            out = np.zeros((info["num_support"], 2), np.float32)
        elif stage.startswith("grid"):
            out = np.zeros((gh, gw, sz.value // 4 // max(1, gw * gh)), np.int32)
        elif stage == "tri0":
            out = np.zeros((info["num_tri0"], 3), np.int32)
        elif stage == "tri1":
            out = np.zeros((info["num_tri1"], 3), np.int32)
        else:
            out = np.zeros((H, W), np.float32)
        if out.nbytes != sz.value:
            raise RuntimeError("stage %s: size %d != expected %d" % (stage, sz.value, out.nbytes))
        if out.nbytes:
            self._check(self._lib.elas_hip_debug_fetch(self._ctx, k, out.ctypes.data, out.nbytes, None))
        return out
