// elas_common.h -- data layout shared by the host engine and the gfx950 kernels.
//
// Every kernel works on a batch of stereo pairs: a device array of ElasPair
// records (one per pair, pointing into the context's HBM arena) and, per
// stage, an inclusive-exclusive prefix of tile counts.  A workgroup walks the
// flattened tile list with a grid stride (a persistent kernel when the grid is
// smaller than the tile count) and finds its pair by binary search on the
// prefix.  A single Elas::process call is a batch of one.
#pragma once
#include <stdint.h>

#define ELAS_MAX_P 16        // prior table entries: |d - d_plane| in [0, 15]
#define ELAS_DENSE_ROWS 3    // image rows per dense tile (a wave = 20 px x 3 rows)
#define ELAS_DENSE_WAVES 8   // waves per dense workgroup; a wave covers dcpw cells per step
#define ELAS_DENSE_THREADS (64 * ELAS_DENSE_WAVES)
#define ELAS_MAX_DISP 511    // disp_max envelope (dense key packs d in 10 bits)
#define ELAS_MAX_KEY_BIAS 16384  // largest -P(k) accepted: cost + bias stays below 2^15 in the dense key

// 16-bit entries of one cell's candidate list (count, largest d, up to
// dmax + 1 disparities), rounded up to 32 bytes
constexpr int elas_cand_stride(int dmax) { return (dmax + 3 + 15) & ~15; }

// Per-triangle raster record (16 floats): plane and span edges, computed once.
//   [0..2] plane a, b, c    [3..8] A_u, A_v, B_u, B_v, C_u, C_v (sorted by u)
//   [9..14] AB_a, AB_b, AC_a, AC_b, BC_a, BC_b         [15] unused
#define ELAS_TRIREC 16

#define ELAS_RT_WS 6         // raster tile: 64 x 32 pixels per workgroup
#define ELAS_RT_HS 5
#define ELAS_RT_W (1 << ELAS_RT_WS)
#define ELAS_RT_H (1 << ELAS_RT_HS)
#define ELAS_ST_MAX 1024     // culling super-tiles per image, at most (LDS masks of k_plane)
#define ELAS_SCAP_MAX 256    // list entries per super-tile, at most (more: the tile scans every chunk)

#define ELAS_DESC_SW 58      // descriptor strip: output columns of a 64-lane wave
#define ELAS_DESC_SH 18      // descriptor strip: output rows by default (18 + 6 rows walked);
                             // a pair's P.dsh is one of 6, 12, 18
#define ELAS_POST_TW 64      // post-filter tile: 64 x 16 output pixels
#define ELAS_POST_TH 16
#define ELAS_GAP_MAX_LINE 15872  // gap interpolation keeps a line in LDS (<= 62 KiB)

// Lines per workgroup of the redundancy filters: a line of len cells takes
// 3 bytes per cell of LDS (int16 value + flag byte).  One line: the
// recurrence takes one wave per line whatever the count, and more
// workgroups in flight beat sharing a workgroup's loads (measured faster
// than two or four lines at 1242x375 and 2816x1880).
constexpr int elas_redun_lines(int len) {
    (void)len;
    return 1;
}
constexpr int elas_redun_lds(int len) {
    return ((2 * elas_redun_lines(len) * len + 15) & ~15) + elas_redun_lines(len) * ((len + 15) & ~15);
}

struct ElasPair {
    int32_t W, H, dmin, dmax;
    int32_t cw, ch, gw, gh;
    int32_t words;           // 32-bit words of a cell's disparity bitmask: (dmax+32)/32
    int32_t cstride;         // 16-bit entries of a cell's candidate list: elas_cand_stride(dmax)
    int32_t step, gsize;     // candidate_stepsize, grid_size
    int32_t dcpw, dspc;      // dense tiling: grid cells per wave per step (1, 2), steps per chunk
    int32_t dsh;             // descriptor strip rows (6, 12 or 18)
    int32_t ntri[2];         // triangles of the left / right mesh (host-known)
    int32_t tri16;           // 1: tri[] holds 16-bit corner indices (fewer than 65536 support points)
    int32_t nblk;            // 256-cell blocks of the support lattice
    int32_t sshift, nsx, nst;  // culling super-tiles: squares of 1 << sshift raster tiles a side,
                               // nsx per row, nst in all
    int32_t scap;            // list entries per super-tile: min(ceil(triangle capacity / 64), ELAS_SCAP_MAX)
    const uint8_t* img[2];   // [H][W]
    uint8_t* desc[2];        // [H][W][16]
    int32_t* can[4];         // [ch][cw]: left pass, LR pass, inconsistency, final
    int32_t* blkcnt;         // [nblk] valid support points per 256-cell block
    float* sp[2];            // [cw*ch+8][2] compacted support points (left, right), then the
                             //   corner points of add_corners
    int32_t* nsp;            // [1]
    float* hsp[2];           // sp[2] again, in mapped pinned host memory (for the host step)
    int32_t* hnsp;           // nsp again, in mapped pinned host memory
    uint32_t* gflag[2];      // [gh*gw][words] scattered disparity flags
    uint16_t* gcand[2];      // [gh*gw][cstride] candidate list of the 3x3-dilated flags: entry 0 the
                             //   count n, entry 1 the largest d (0 when n = 0), entries 2.. the n
                             //   disparities ascending; border cells n = 0.  gcand[1] follows
                             //   gcand[0] in one block
    int32_t* glist[2];       // [gh*gw][dmax+2] reference list layout (debug only)
    const int32_t* tri[2];   // [ntri][3] int32, or uint16 when tri16 (mapped pinned memory)
    float* trirec[2];        // [ntri][ELAS_TRIREC]: (A_u, B_u, C_u, box rows v0 | v1 << 16), AB and AC
                             //   edges, BC edge, plane (a, b, c); the box is half-open, rows 0
                             //   when empty
    int32_t* scnt[2];        // [nst] entries appended to a super-tile's list (zeroed in phase A);
                             //       above scap the list is incomplete
    int32_t* slist[2];       // [nst][scap][4] per super-tile, in no particular order: chunk c and the
                             //   64-bit mask (lo, hi) of its triangles 64c+k that meet the super-tile
    int32_t* prior[2];       // [H][W] (int)plane(u,v) of the covering triangle
                             //        with the highest index, -1 if uncovered
    float* mean[2];          // [H][W] (float)prior: the reference's map (debug only)
    float* draw[2];          // [H][W] dense disparity before the LR check
    float* dout[2];          // [H][W] final disparity
    int32_t* pbuf[2];        // [2][H][W] post-filter scratch (labels, sizes / ping-pong map)
};

// Uniform parameters of a launch (from elas_params).
struct ElasKParams {
    float support_threshold;
    int32_t support_texture;
    int32_t incon_window, incon_threshold, incon_min_support;
    int32_t lr_threshold;
    int32_t plane_radius;
    int32_t match_texture;
    int32_t key_bias;        // added to dense costs so keys stay non-negative
    int32_t all_p_negative;  // every prior P(k), k <= plane_radius, is < 0
    int32_t P[ELAS_MAX_P];
    // post-processing (off unless enabled)
    float speckle_sim;
    int32_t speckle_size, ipol_gap, adaptive_mean, median, post_images;  // post_images: 1 (left) or 2
    int32_t add_corners;     // corner support points after compaction
};

// Launch a stage over a batch: pairs/prefix live in device memory, prefix has
// npairs+1 entries, total = prefix[npairs].  Both are read-only for the
// duration of a launch; device code reads them through the constant address
// space, so they come in through scalar loads that do not wait on the
// kernel's own vector loads and stores (vector loads of them would have to).
#if defined(__HIP_DEVICE_COMPILE__)
#define ELAS_CONST_AS __attribute__((address_space(4)))
#else
#define ELAS_CONST_AS
#endif
struct ElasLaunch {
    const ELAS_CONST_AS ElasPair* pairs;
    const ELAS_CONST_AS int32_t* prefix;
    int32_t npairs;
    int32_t total;
};

// Stage ids (one prefix table per stage).  S_DENSE holds the dense units of
// the pairs tiled with one cell per wave, S_DENSE2 those with two; S_DESC,
// S_DESC12 and S_DESC6 the descriptor strips of the pairs with 18-, 12- and
// 6-row strips.  S_PLANE counts 64-triangle chunks of both meshes, S_RTILES
// the raster tiles of both images, S_PAIRS one tile per pair.
enum ElasStageId {
    S_DESC = 0, S_SUPPORT, S_INCON, S_REDUN_H, S_REDUN_V, S_BLOCKS, S_GRID_CELLS,
    S_PLANE, S_RTILES, S_DENSE, S_PIXELS, S_GRID_FILL, S_POST_TILES, S_POST_ROWS, S_POST_COLS,
    S_DENSE2, S_DESC12, S_DESC6, S_PAIRS, S_NUM_STAGES
};
