// elas_launch.h -- host entry points of the gfx950 kernels (defined in the
// .hip files, called by the host engine).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include "elas_common.h"

// Scratch of the entry point that allocates nothing: ELAS_SCRATCH_BYTES of
// device memory per GPU (a file-scope device array).  elas_scratch_setup()
// returns its address on `device`, which must be the current device; the
// first call per device looks the address up (no allocation), later calls
// return the kept address.  0 on success.
#define ELAS_SCRATCH_BYTES ((size_t)128 << 20)
#define ELAS_SCRATCH_DEVICES 64
int elas_scratch_setup(int device, void** ptr);

void elas_launch_descriptor(const ElasLaunch& L, int sh, int grid, hipStream_t s);
// Support matching, left pass and right-to-left check in one launch (lds:
// the right pass's staging, the larger of the two).
void elas_launch_support(const ElasLaunch& L, const ElasKParams& kp, int grid, size_t lds, hipStream_t s);
void elas_launch_incon(const ElasLaunch& L, const ElasKParams& kp, int grid, hipStream_t s);
void elas_launch_redun(const ElasLaunch& L, bool vertical, int grid, size_t lds, hipStream_t s);
void elas_launch_blkcount(const ElasLaunch& L, int grid, hipStream_t s);
void elas_launch_compact(const ElasLaunch& L, int grid, hipStream_t s);
// Corner support points of add_corners, appended after compaction (one
// workgroup per pair; S_PAIRS tiles).
void elas_launch_corners(const ElasLaunch& L, int grid, hipStream_t s);
void elas_launch_grid_fill(const ElasLaunch& L, int grid, hipStream_t s);
void elas_launch_grid_scatter(const ElasLaunch& L, int grid, hipStream_t s);
void elas_launch_grid_dilate(const ElasLaunch& L, int grid, hipStream_t s);
void elas_launch_grid_list(const ElasLaunch& L, int grid, hipStream_t s);
void elas_launch_plane(const ElasLaunch& L, int grid, size_t lds, hipStream_t s);
void elas_launch_rtile(const ElasLaunch& L, int grid, hipStream_t s);
// Polls flag (mapped pinned memory) until it is non-zero, for at most `limit`
// ticks of the 100 MHz device clock, and copies `words` 32-bit words to dst:
// from `ready` when the flag was set, else from `fallback` (then status[0] = 1).
void elas_launch_gate(const uint32_t* flag, uint32_t* status, const uint32_t* ready, const uint32_t* fallback,
                      uint32_t* dst, int words, uint64_t limit, hipStream_t s);
void elas_launch_mean(const ElasLaunch& L, int grid, hipStream_t s);
void elas_launch_dense(const ElasLaunch& L, const ElasKParams& kp, int cpw, int grid, size_t lds, hipStream_t s);
// Lets the dense kernels allocate up to `bytes` of dynamic LDS; 0 on success.
int elas_dense_allow_lds(size_t bytes);
void elas_launch_lr(const ElasLaunch& L, const ElasKParams& kp, int grid, hipStream_t s);
void elas_launch_speckle(const ElasLaunch& Ltiles, const ElasLaunch& Lpix, const ElasKParams& kp, int grid,
                         hipStream_t s);
void elas_launch_gap(const ElasLaunch& Lrows, const ElasLaunch& Lcols, const ElasKParams& kp, int grid,
                     size_t lds_rows, size_t lds_cols, hipStream_t s);
void elas_launch_sep(const ElasLaunch& Ltiles, const ElasKParams& kp, bool median, int from_tmp, int grid,
                     hipStream_t s);
void elas_launch_post_copy(const ElasLaunch& Lpix, const ElasKParams& kp, int grid, hipStream_t s);
