/*
 * elas_hip.h -- C ABI of the MI355X-native ELAS stereo path.
 *
 * Plain pointers and sizes only: a cgo / ctypes / JNI binding needs nothing
 * else.  Every entry point returns an elas_status (0 = success); the message
 * of the last failure on a context is available from elas_hip_last_error().
 * The reference has no return codes (CUDA errors and Triangle's "fewer than
 * three points" call exit(), src/Utils.h:11-38, src/triangle.cpp:7652-7654);
 * here a non-zero status replaces every exit().
 *
 * Reference interface replaced (paths relative to the reference repository):
 *   Elas::parameters                 src/Elas.h:14-67
 *   Elas::Elas()                     src/Elas.h:70-72
 *   Elas::process(l, r, D1, D2, dims) src/Elas.h:76-83, src/Elas.cu:19-120
 */
#ifndef ELAS_HIP_H
#define ELAS_HIP_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Same fields, types and order as Elas::parameters (src/Elas.h:15-37), so a
 * pointer to the reference struct can be passed unchanged.  The values the
 * reference kernels actually use are compile-time macros
 * (src/Kernels/KernelsCommon.h:8-27); elas_params_default() returns those
 * values, with every post-processing switch off because the reference
 * pipeline ends at the left/right check (src/Elas.cu:107-119). */
typedef struct elas_params {
    int32_t disp_min;              /* min disparity                              */
    int32_t disp_max;              /* max disparity                              */
    float   support_threshold;     /* max. uniqueness ratio (best vs. 2nd best)  */
    int32_t support_texture;       /* min texture for support points             */
    int32_t candidate_stepsize;    /* step of the support-point lattice          */
    int32_t incon_window_size;     /* half window of the inconsistency check     */
    int32_t incon_threshold;       /* disparity similarity for consistency       */
    int32_t incon_min_support;     /* min number of consistent support points    */
    bool    add_corners;           /* corner support points (see below)          */
    int32_t grid_size;             /* cell size of the disparity grid            */
    float   beta;                  /* image likelihood parameter                 */
    float   gamma;                 /* prior constant                             */
    float   sigma;                 /* prior sigma                                */
    float   sradius;               /* prior sigma radius                         */
    int32_t match_texture;         /* min texture for dense matching             */
    int32_t lr_threshold;          /* left/right consistency threshold           */
    float   speckle_sim_threshold; /* speckle segmentation similarity            */
    int32_t speckle_size;          /* max speckle size removed (0 = off)         */
    int32_t ipol_gap_width;        /* interpolated gap width (0 = off)           */
    bool    filter_median;         /* optional median filter                     */
    bool    filter_adaptive_mean;  /* optional adaptive mean filter              */
    bool    postprocess_only_left; /* post-process D1 only                       */
    bool    subsampling;           /* refused: ELAS_ERR_UNSUPPORTED (see below)  */
} elas_params;

/* Two switches the reference declares but never implements (src/Elas.h:23,
 * 36-38):
 *   add_corners  after compaction, each image corner (0,0), (0,H-1), (W-1,0),
 *                (W-1,H-1) becomes a support point with the disparity of the
 *                nearest support point (smallest squared distance, the first in
 *                compaction order on ties), and the two right corners are
 *                added again at u = W-1+d so the right image's mesh reaches its
 *                right border: 6 points after the nsp lattice points (only when
 *                nsp >= 1).  The disparity grid stays the lattice's.  This is
 *                this project's reading of the field comment; the oracle
 *                defines it the same way.
 *   subsampling  (D1/D2 of W/2 x H/2) is refused with ELAS_ERR_UNSUPPORTED by
 *                create and set_params: the reference gives no semantics for
 *                which pixels are matched or for the left/right check at half
 *                resolution. */
typedef enum elas_status {
    ELAS_OK = 0,
    ELAS_ERR_INVALID_ARGUMENT = 1, /* bad dims, null pointer, bad parameters     */
    ELAS_ERR_HIP = 2,              /* a HIP runtime call or launch failed        */
    ELAS_ERR_OUT_OF_MEMORY = 3,    /* device or host allocation failed           */
    ELAS_ERR_TOO_FEW_POINTS = 4,   /* < 3 support points: D1/D2 are all -1       */
    ELAS_ERR_NO_DEVICE = 5,        /* no usable gfx950 device                    */
    ELAS_ERR_UNSUPPORTED = 6,      /* request outside the supported envelope     */
    ELAS_ERR_BUSY = 7,             /* serving queue full / budget exhausted      */
    ELAS_ERR_IO = 8                /* unreadable, malformed or unwritable file   */
} elas_status;

/* Intermediates kept by the last process call, for parity dumps
 * (elas_hip_debug_fetch).  Layouts follow the reference buffers. */
typedef enum elas_stage {
    ELAS_STAGE_DESC0 = 0,      /* u8  [H][W][16] left descriptor (0 where unwritten)  */
    ELAS_STAGE_DESC1 = 1,      /* u8  [H][W][16] right descriptor                     */
    ELAS_STAGE_CAN_LEFT = 2,   /* i32 [ch][cw] after the left->right support pass     */
    ELAS_STAGE_CAN_LR = 3,     /* i32 [ch][cw] after the right->left pass (LR test)   */
    ELAS_STAGE_CAN_INCON = 4,  /* i32 [ch][cw] after the inconsistency filter         */
    ELAS_STAGE_CAN_FINAL = 5,  /* i32 [ch][cw] after the redundancy filters           */
    ELAS_STAGE_SP0 = 6,        /* f32 [nsp][2] left support points (u, v)             */
    ELAS_STAGE_SP1 = 7,        /* f32 [nsp][2] right support points (u - d, v)        */
    ELAS_STAGE_GRID0 = 8,      /* i32 [gh][gw][disp_max+2] left disparity grid        */
    ELAS_STAGE_GRID1 = 9,      /* i32 [gh][gw][disp_max+2] right disparity grid       */
    ELAS_STAGE_TRI0 = 10,      /* i32 [ntri0][3] left Delaunay triangles              */
    ELAS_STAGE_TRI1 = 11,      /* i32 [ntri1][3] right Delaunay triangles             */
    ELAS_STAGE_MEAN0 = 12,     /* f32 [H][W] left plane prior (-1 = uncovered)        */
    ELAS_STAGE_MEAN1 = 13,     /* f32 [H][W] right plane prior                        */
    ELAS_STAGE_D1_RAW = 14,    /* f32 [H][W] left dense disparity before the LR check */
    ELAS_STAGE_D2_RAW = 15,    /* f32 [H][W] right dense disparity before the LR check*/
    ELAS_STAGE_D1 = 16,        /* f32 [H][W] final left disparity                     */
    ELAS_STAGE_D2 = 17,        /* f32 [H][W] final right disparity                    */
    ELAS_STAGE_COUNT = 18
} elas_stage;

/* Sizes of the last processed pair (valid after a successful process call). */
typedef struct elas_shape_info {
    int32_t width, height;        /* image                                  */
    int32_t can_width, can_height;/* support lattice: W/step, H/step        */
    int32_t grid_width, grid_height; /* disparity grid: ceil(W/gs), ceil(H/gs) */
    int32_t num_support;          /* support points after compaction        */
    int32_t num_tri0, num_tri1;   /* Delaunay triangles left / right        */
    int32_t status;               /* elas_status of that call               */
} elas_shape_info;

/* Per-stage device time of the last call in milliseconds (HIP events).  The
 * stages inside phases A and B read -1 when the phases ran as graphs
 * (elas_hip_set_graphs); PHASE_A / PHASE_B cover them in either mode. */
typedef enum elas_timer {
    ELAS_T_DESCRIPTOR = 0, ELAS_T_SUPPORT = 1, ELAS_T_FILTER = 2, ELAS_T_COMPACT = 3,
    ELAS_T_GRID = 4, ELAS_T_TRIANGLE = 5, ELAS_T_PLANE = 6, ELAS_T_DENSE = 7,
    ELAS_T_LR = 8, ELAS_T_POST = 9, ELAS_T_TOTAL = 10, ELAS_T_PHASE_A = 11,
    ELAS_T_PHASE_B = 12,
    /* host step (wall clock): support points read (the compaction kernel
     * writes them into pinned memory), Triangle jobs from submit to the last
     * done, the longest job */
    ELAS_T_SP_DOWNLOAD = 13, ELAS_T_TRI_WALL = 14, ELAS_T_TRI_JOB = 15,
    /* host wall clock from the call's start to phase A launched, and from the
     * meshes done to phase B launched; device time from phase A's start to the
     * last copy of the call, and the device's idle gap between the phases */
    ELAS_T_HOST_A = 16, ELAS_T_HOST_B = 17, ELAS_T_GPU_SPAN = 18, ELAS_T_AB_GAP = 19,
    ELAS_T_COUNT = 20
} elas_timer;

typedef struct elas_ctx elas_ctx;

/* Effective reference defaults (src/Kernels/KernelsCommon.h:8-27). */
void elas_params_default(elas_params* p);
/* The dead initialisers of src/Elas.h:42-66 (post-processing on). */
void elas_params_declared(elas_params* p);

/* Create a context on HIP device `device` (replaces Elas::Elas(),
 * src/Elas.h:70-72).  Device buffers are pooled in the context and grow on
 * demand; nothing is allocated per call once a shape has been seen. */
int elas_hip_create(const elas_params* params, int device, elas_ctx** out);
int elas_hip_destroy(elas_ctx* ctx);
int elas_hip_set_params(elas_ctx* ctx, const elas_params* params);
const char* elas_hip_last_error(const elas_ctx* ctx);
/* HIP stream the context enqueues on (a hipStream_t). */
void* elas_hip_stream(elas_ctx* ctx);

/* Elas::process (src/Elas.h:76-83, src/Elas.cu:19-120): host buffers in and
 * out, synchronous.  dims[0] = width, dims[1] = height (dims[2], the
 * bytes-per-line of upstream libelas, is not read).  Images are W*H u8
 * row-major; D1/D2 are W*H f32 with -1 for invalid pixels. */
int elas_hip_process(elas_ctx* ctx, const uint8_t* left, const uint8_t* right,
                     float* D1, float* D2, const int32_t* dims);

/* Same computation on device-resident buffers (hipMalloc'd by the caller):
 * no H2D/D2H of images or disparities.  Synchronous. */
int elas_hip_process_device(elas_ctx* ctx, const uint8_t* d_left, const uint8_t* d_right,
                            float* d_D1, float* d_D2, int32_t width, int32_t height);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): elas_hip_process_device without any device or pinned
 * allocation.  Its intermediates live in a per-GPU scratch array of 128 MiB
 * (a file-scope device array of the library, held by every GPU the library
 * runs on until the process ends) and in pinned staging and streams made when
 * the context was created; phases run as direct launches, and no autotuning
 * happens (a shape uses the tiling an earlier call on this context tuned, else
 * the built-in heuristic).  The scratch bounds the pair: about 1.5 Mpx at
 * disparities 0..255 with post-processing off (1242x375, 1280x960, 1920x720);
 * a larger pair returns ELAS_ERR_UNSUPPORTED before any work.  All contexts
 * on one GPU share its scratch, which keeps the last call's contents: calls
 * that overlap on one GPU (two threads, two contexts) race and must be
 * serialised by the caller.  Synchronous; same results as
 * elas_hip_process_device.  The host Delaunay step still uses the host heap. */
int elas_hip_process_device_noalloc(elas_ctx* ctx, const uint8_t* d_left, const uint8_t* d_right,
                                    float* d_D1, float* d_D2, int32_t width, int32_t height);

/* Batched form: n pairs of possibly different sizes and disparity ranges,
 * each stage one launch over all pairs.  Arrays of n host pointers; dims is
 * n*3 int32: (W, H, disp_max), disp_max < 0 meaning the context's.
 * statuses (optional) receives one elas_status per pair; the call returns
 * ELAS_OK when every pair either succeeded or had too few support points. */
int elas_hip_process_batch(elas_ctx* ctx, int32_t n, const uint8_t* const* left,
                           const uint8_t* const* right, float* const* D1, float* const* D2,
                           const int32_t* dims, int32_t* statuses);

/* Batched form on device-resident buffers (arrays of n device pointers). */
int elas_hip_process_batch_device(elas_ctx* ctx, int32_t n, const uint8_t* const* d_left,
                                  const uint8_t* const* d_right, float* const* d_D1,
                                  float* const* d_D2, const int32_t* dims, int32_t* statuses);

/* Launch mode (no reference counterpart): phases A (descriptor .. support
 * point count) and B (plane .. post-processing) are captured as HIP graphs
 * once per shape bucket and replayed (default), or launched kernel by kernel
 * with per-stage timing events (on = 0).  Results are identical. */
int elas_hip_set_graphs(elas_ctx* ctx, int32_t on);

/* Pipelining of the batched forms: a batch is cut into chunks of up to
 * `chunk` pairs (0: the whole batch is one chunk) that move through `slots`
 * pipeline slots (1..8), so the GPU stages of one chunk overlap the host
 * Delaunay step of another.  Each slot keeps its own device arena sized for
 * the largest chunk.  Defaults: 3 slots, chunk 0.  No reference counterpart
 * (the reference processes one pair per call). */
int elas_hip_set_pipeline(elas_ctx* ctx, int32_t slots, int32_t chunk);

/* Intermediates of the last process call on this context (kept only when
 * elas_hip_set_debug(ctx, 1)).  Copies min(bytes, stage size) bytes to dst
 * and stores the stage size in *stage_bytes (either may be NULL/0 to query).
 * flags & 2 (a test of the raster's overflow path) caps the raster's
 * per-tile triangle lists at one entry, so every tile with triangles from
 * more than one 64-triangle chunk tests the whole mesh; results are the same.
 * flags & 4 turns off the phase B gate of single-pair calls (phase B is then
 * launched after the host Delaunay step instead of being queued during it
 * behind a kernel that waits for the host's release); results are the same.
 * flags & 8 (a test of the gate's failure path) bounds the gate's wait at
 * one tick instead of 30 s, so a gated single-pair call returns ELAS_ERR_HIP
 * ("phase B gate timed out") with its maps unspecified; the next call
 * without the flag is unaffected. */
int elas_hip_set_debug(elas_ctx* ctx, int flags);
int elas_hip_debug_fetch(elas_ctx* ctx, int stage, void* dst, size_t bytes, size_t* stage_bytes);
int elas_hip_shape_info(const elas_ctx* ctx, elas_shape_info* info);
/* Per-stage device milliseconds of the last call (ELAS_T_COUNT floats). */
int elas_hip_stage_times(const elas_ctx* ctx, float* ms, int32_t n);

/* Kernel timing (no reference counterpart): re-launches one kernel of the
 * last call `reps` times back to back on the context's stream and returns the
 * average milliseconds per launch between HIP events.  Every kernel listed
 * here recomputes the same outputs from the same inputs, so the results of
 * the last call are unchanged. */
typedef enum elas_kernel {
    ELAS_K_DESCRIPTOR = 0, ELAS_K_SUPPORT = 1, /* both support passes, one launch */
    ELAS_K_DENSE = 2, ELAS_K_LR = 3, ELAS_K_COUNT = 4
} elas_kernel;
int elas_hip_time_kernel(elas_ctx* ctx, int32_t kernel, int32_t reps, float* avg_ms);

/* Dense-matching tile autotuner (no reference counterpart: the reference
 * runs one block per image row, KerComputeDisparity.cu:156-316).  The dense
 * kernel works on bands of image rows cut into chunks of steps; a step is
 * 8 waves x cells_per_wave grid cells wide.  With autotune on (default), the
 * first successful call with a new (W, H, disp_max) times every candidate
 * tiling on that pair after its results are complete and keeps the fastest
 * for later calls (results are identical under every tiling).  Off: a
 * built-in heuristic.  elas_hip_set_dense_tiles forces a tiling
 * (cells_per_wave 1 or 2, steps_per_chunk >= 1 or 0 for the heuristic's;
 * 0, 0 restores automatic tiling).  elas_hip_dense_tiles reports the tiling
 * a shape would use now and its tuned microseconds per launch (-1 if not
 * tuned). */
int elas_hip_set_autotune(elas_ctx* ctx, int32_t on);
int elas_hip_set_dense_tiles(elas_ctx* ctx, int32_t cells_per_wave, int32_t steps_per_chunk);
int elas_hip_dense_tiles(elas_ctx* ctx, int32_t W, int32_t H, int32_t disp_max, int32_t* cells_per_wave,
                         int32_t* steps_per_chunk, float* tuned_us);
/* The same for the descriptor tile (replaces the fixed 32x32 blocks of
 * src/Elas.cu:131-154): a wave computes a strip of 58 columns x rows rows
 * (6, 12 or 18; short strips give a small image more waves).  The autotuner
 * times the three on the first pair of a shape; elas_hip_set_desc_rows
 * forces one (0: automatic). */
int elas_hip_set_desc_rows(elas_ctx* ctx, int32_t rows);
int elas_hip_desc_rows(elas_ctx* ctx, int32_t W, int32_t H, int32_t disp_max, int32_t* rows, float* tuned_us);

/* ---- Caller-side image I/O (host only; no device needed) ------------------
 * elas_io_read_pgm replaces loadPGM (src/Image.h:129-160): binary "P5" with
 * '#' comment lines between header tokens and maxval <= 255; dst == NULL
 * only reports width/height; a malformed or truncated file returns
 * ELAS_ERR_IO (the reference throws on a bad magic or maxval and reads short
 * files silently).  elas_io_write_pgm replaces savePGM (src/Image.h:162-169).
 * elas_io_write_pfm writes a float map as PFM ("Pf", scale -1, bottom row
 * first).  elas_io_disparity_to_u8 is the 8-bit view the reference displays
 * (src/Main.cu:14-27): (uint8)max(255.0 * d / max(d), 0.0); all zeros when no
 * value is positive (the reference divides by zero there). */
int elas_io_read_pgm(const char* path, uint8_t* dst, int64_t cap, int32_t* width, int32_t* height);
int elas_io_write_pgm(const char* path, const uint8_t* src, int32_t width, int32_t height);
int elas_io_write_pfm(const char* path, const float* src, int32_t width, int32_t height);
int elas_io_disparity_to_u8(const float* d, uint8_t* out, int64_t n);

/* Host Delaunay triangulation as the pipeline does it: the reference's
 * Triangle with switches "zQB" on n points [x0, y0, x1, y1, ...]
 * (replaces Elas::computeDelaunyTrianguation, src/Elas.cu:278-328, for tests
 * and host tools).  Writes min(count, cap) triangles of three zero-based point
 * indices to tri and the count to *ntri.  Needs no device.  Fewer than 3 points
 * returns ELAS_ERR_TOO_FEW_POINTS (the reference's Triangle exits there). */
int elas_hip_triangulate(const float* pts, int32_t n, int32_t* tri, int64_t cap, int64_t* ntri);

/* Threads of the host Delaunay step (no reference counterpart: the reference
 * runs Triangle serially).  Triangle's divide-and-conquer allocates its
 * triangles in post-order (left half, right half, the merge of the two), so
 * its top max_depth levels can run as up to 2^max_depth concurrent subtrees,
 * each with its own triangle pool, merged by Triangle's own mergehulls() and
 * read out in that order: the triangle list, numbering included, is the
 * serial one under every setting.  Process-wide; 0 = serial Triangle, -1 =
 * automatic (the default: subtrees of at least 128 points, at most 8 per
 * mesh).  elas_hip_tri_split_depth reports the depth a mesh of n points gets
 * under the current setting. */
int elas_hip_set_tri_split(int32_t max_depth);
int elas_hip_tri_split_depth(int32_t n);

/* ---- Serving (no reference counterpart: the reference has one blocking
 * process call).  A scheduler thread owned by the context admits queued
 * requests into idle pipeline slots as chunks, FIFO, as long as the chunk's
 * device footprint fits one slot's share of the HBM budget; requests that
 * arrive while others are on the GPU join the next chunk, so the GPU stages
 * of several chunks and the host Delaunay step of another overlap.  While
 * serving, the blocking calls on this context return ELAS_ERR_BUSY. */
typedef struct elas_serve_config {
    uint64_t hbm_budget_bytes; /* device bytes for all slot arenas (0: 8 GiB)  */
    int32_t max_chunk_pairs;   /* cap on pairs per chunk (0: 32)                */
    int32_t slots;             /* pipeline slots (0: the context's, default 3)  */
} elas_serve_config;

typedef struct elas_serve_stats {
    int64_t submitted, completed, chunks;
    int64_t max_chunk_pairs;       /* largest chunk formed                     */
    int64_t max_in_flight_pairs;   /* most pairs on the device at once          */
    uint64_t slot_capacity_bytes;  /* arena bytes per slot (budget / slots)     */
    uint64_t peak_in_flight_bytes; /* largest sum of admitted chunk footprints  */
} elas_serve_stats;

int elas_hip_serve_start(elas_ctx* ctx, const elas_serve_config* cfg);
/* Queue one pair.  Host buffers unless flags & 1 (device-resident buffers).
 * Buffers must stay valid until the request completes; pinned host buffers
 * (hipHostMalloc) let the copies overlap the GPU work.  disp_max < 0: the
 * context's.  Returns ELAS_ERR_OUT_OF_MEMORY at once for a pair larger than
 * a slot's share of the budget. */
int elas_hip_submit(elas_ctx* ctx, const uint8_t* left, const uint8_t* right, float* D1, float* D2,
                    int32_t width, int32_t height, int32_t disp_max, int32_t flags, int64_t* ticket);
/* ELAS_OK once the request has finished (its status in *status, its
 * submit-to-completion time in *latency_ms; the ticket is then retired),
 * ELAS_ERR_BUSY while it is queued or in flight. */
int elas_hip_poll(elas_ctx* ctx, int64_t ticket, int32_t* status, float* latency_ms);
/* Waits for every queued request, then stops the scheduler thread. */
int elas_hip_serve_stop(elas_ctx* ctx);
int elas_hip_serve_stats(elas_ctx* ctx, elas_serve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* ELAS_HIP_H */
