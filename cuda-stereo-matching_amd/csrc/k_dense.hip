// k_dense.hip -- plane prior, triangle rasterisation, dense matching and the
// left/right check (gfx950).  Compiled with -ffp-contract=off; the float
// plane and span arithmetic also uses explicit _rn intrinsics so no FMA is
// formed.
#include <hip/hip_runtime.h>
#include "elas_common.h"
#include "elas_device.h"

// ---- plane solve: kerSolve, KerComputeDisparity.cu:322-375 ---------------
// Gauss-Jordan with full pivoting in double (eps 1e-20, ties on >= keep the
// last pivot), the solution overwriting B.
__device__ bool elas_solve3(double A[3][3], double B[3]) {
    const double eps = 1e-20;
    int ipiv[3] = {0, 0, 0};
    int irow = 0, icol = 0;
    for (int i = 0; i < 3; i++) {
        double big = 0.0;
        for (int j = 0; j < 3; j++)
            if (ipiv[j] != 1)
                for (int k = 0; k < 3; k++)
                    if (ipiv[k] == 0)
                        if (fabs(A[j][k]) >= big) { big = fabs(A[j][k]); irow = j; icol = k; }
        ++(ipiv[icol]);
        if (irow != icol) {
            for (int l = 0; l < 3; l++) { const double tmp = A[irow][l]; A[irow][l] = A[icol][l]; A[icol][l] = tmp; }
            const double tb = B[irow]; B[irow] = B[icol]; B[icol] = tb;
        }
        if (fabs(A[icol][icol]) < eps) return false;
        const double pivinv = 1.0 / A[icol][icol];
        A[icol][icol] = 1.0;
        for (int l = 0; l < 3; l++) A[icol][l] = __dmul_rn(A[icol][l], pivinv);
        B[icol] = __dmul_rn(B[icol], pivinv);
        for (int ll = 0; ll < 3; ll++)
            if (ll != icol) {
                const double dum = A[ll][icol];
                A[ll][icol] = 0.0;
                for (int l = 0; l < 3; l++) A[ll][l] = __dsub_rn(A[ll][l], __dmul_rn(A[icol][l], dum));
                B[ll] = __dsub_rn(B[ll], __dmul_rn(B[icol], dum));
            }
    }
    return true;
}

// ---- raster bins ------------------------------------------------------------
// Tiles (ELAS_RASTER_T px) a triangle's covered pixels can fall in.  Covered
// pixels lie in columns [min u, max u) (the reference's spans are half-open
// between corner columns, KerComputeDisparity.cu:128-150) and in rows that
// are truncated float evaluations of the edge lines, within the corner rows
// up to rounding: with slopes s, |u| <= U and rows <= V the float error of
// s*u + (v0 - s*u0) is below 8 * 2^-24 * (3 s U + 3 V) rows.  The rows are
// widened by 3 + that bound (rounded up with room for the float evaluation
// of the bound itself), so the tile range is a superset; rasterisation
// tests every pixel exactly.
__device__ __forceinline__ bool elas_bin_box(const float* rec, int W, int H, int& tx0, int& tx1, int& ty0,
                                             int& ty1) {
    const int cu0 = (int)rec[3], cv0 = (int)rec[4], cu1 = (int)rec[5], cv1 = (int)rec[6];
    const int cu2 = (int)rec[7], cv2 = (int)rec[8];  // corners sorted by u
    const int vmin = min(cv0, min(cv1, cv2)), vmax = max(cv0, max(cv1, cv2));
    float smax = 0.0f;
    if (cu0 != cu1) smax = fmaxf(smax, fabsf((float)(cv0 - cv1) / (float)(cu0 - cu1)));
    if (cu0 != cu2) smax = fmaxf(smax, fabsf((float)(cv0 - cv2) / (float)(cu0 - cu2)));
    if (cu1 != cu2) smax = fmaxf(smax, fabsf((float)(cv1 - cv2) / (float)(cu1 - cu2)));
    const int Ui = max(abs(cu0), abs(cu2)) + 1, Vi = max(abs(vmin), abs(vmax)) + 1;
    const float xf = 8.0f * (3.0f * smax * (float)Ui + 3.0f * (float)Vi) * (1.0f / 16777216.0f);
    const int margin = xf < 100.0f ? 3 + (int)(xf + 1e-3f) : 4 + (int)(xf * 1.001f);
    const int u0 = max(cu0, 0), u1 = min(cu2, W);
    const int v0 = max(vmin - margin, 0), v1 = min(vmax + margin + 1, H);
    if (u0 >= u1 || v0 >= v1) return false;
    tx0 = u0 / ELAS_RASTER_T;
    tx1 = (u1 - 1) / ELAS_RASTER_T;
    ty0 = v0 / ELAS_RASTER_T;
    ty1 = (v1 - 1) / ELAS_RASTER_T;
    return true;
}

// The (triangle, tile) items of a 64-lane workgroup's triangles, spread evenly
// over the lanes: a triangle's box is a few tiles on average but up to a few
// hundred for slivers along the hull, and a lane walking its own box would
// hold the whole wave.  Lane i's triangle owns items [incl[i-1], incl[i]); item
// k is tile (ty0 + k / w, tx0 + k % w) of its box.  Calls f(tag, tile row,
// tile column) per item.  s holds 5 * 64 ints of LDS.
template <typename F>
__device__ __forceinline__ void elas_bin_items(int* s, bool ok, int tx0, int tx1, int ty0, int ty1, int tag, F f) {
    const int lane = threadIdx.x;
    const int w = ok ? tx1 - tx0 + 1 : 1;
    const int nt = ok ? w * (ty1 - ty0 + 1) : 0;
    int incl = nt;
    for (int o = 1; o < 64; o <<= 1) {
        const int x = __shfl_up(incl, o);
        if (lane >= o) incl += x;
    }
    const int total = __shfl(incl, 63);
    int* s_incl = s;
    int* s_tx = s + 64;
    int* s_ty = s + 128;
    int* s_w = s + 192;
    int* s_tag = s + 256;
    s_incl[lane] = incl; s_tx[lane] = tx0; s_ty[lane] = ty0; s_w[lane] = w; s_tag[lane] = tag;
    elas_lds_barrier();
    for (int item = lane; item < total; item += 64) {
        // owner: the number of lanes whose inclusive count is <= item (lane 63's
        // count, the total, is above every item)
        int j = 0;
        #pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if (s_incl[j + step - 1] <= item) j += step;
        const int k = item - (j ? s_incl[j - 1] : 0);
        const int ww = s_w[j];
        const int q = k / ww;
        f(s_tag[j], s_ty[j] + q, s_tx[j] + (k - q * ww));
    }
    elas_lds_barrier();  // the arrays are rewritten by the next grid-stride step
}

// Clears the per-tile counts of both images and the overflow flag.
__global__ __launch_bounds__(256) void k_bin_clear(ElasLaunch L) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        if (!P.gpu_bins) continue;
        const int nt1 = P.ntx * P.nty + 1;
        const int i = (t - L.prefix[p]) * 256 + threadIdx.x;
        if (i < nt1) elas_st(P.bcnt[0] + i, 0);
        else if (i < 2 * nt1) elas_st(P.bcnt[1] + i - nt1, 0);
        else if (i == 2 * nt1) elas_st(P.bovf, 0);
    }
}

// Exclusive scan of one image's tile counts into list offsets (and fill
// cursors); one workgroup per image.  Lists that would outgrow the capacity
// set the overflow flag and are left empty.
__global__ __launch_bounds__(1024) void k_bin_scan(ElasLaunch L) {
    __shared__ int s_sum[1024];
    const int tid = threadIdx.x;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        if (!P.gpu_bins) continue;
        const int img = t - L.prefix[p];
        const int n = P.ntx * P.nty;
        int32_t* cnt = P.bcnt[img];
        int32_t* cur = P.bcur[img];
        const int C = (n + 1023) / 1024;
        const int a = min(n, tid * C), b = min(n, a + C);
        int sum = 0;
        for (int i = a; i < b; i++) sum += cnt[i];
        s_sum[tid] = sum;
        elas_lds_barrier();
        for (int off = 1; off < 1024; off <<= 1) {
            const int x = tid >= off ? s_sum[tid - off] : 0;
            elas_lds_barrier();
            s_sum[tid] += x;
            elas_lds_barrier();
        }
        const int total = s_sum[1023];
        const bool over = total > P.bcap;
        int run = tid > 0 ? s_sum[tid - 1] : 0;
        for (int i = a; i < b; i++) {
            const int c = cnt[i];
            cnt[i] = over ? 0 : run;
            cur[i] = over ? 0 : run;
            run += c;
        }
        if (tid == 0) {
            cnt[n] = over ? 0 : total;
            if (over) elas_st(P.bovf, 1);
        }
        elas_lds_barrier();
    }
}

// Fills the tile lists from the plane records (which hold the sorted corners),
// a workgroup's (triangle, tile) items spread over its lanes; the order within
// a list does not matter to the raster.
__global__ __launch_bounds__(64) void k_bin_fill(ElasLaunch L) {
    __shared__ int s_items[5 * 64];
    const int total = L.prefix[L.npairs];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        if (!P.gpu_bins) continue;
        const int idx = (t - L.prefix[p]) * 64 + threadIdx.x;
        const int n0 = P.ntri[0], n1 = P.ntri[1];
        // an image whose lists overflowed has all offsets 0 (k_bin_scan): its
        // writes below stay within the capacity and are never read
        const bool live = idx < n0 + n1;
        const int img = live && idx >= n0;
        const int ti = img ? idx - n0 : idx;
        int tx0 = 0, tx1 = 0, ty0 = 0, ty1 = 0;
        const bool ok = live && elas_bin_box(P.trirec[img] + (size_t)ti * ELAS_TRIREC, P.W, P.H, tx0, tx1, ty0, ty1);
        const int cap = P.bcap, ntx = P.ntx;
        elas_bin_items(s_items, ok, tx0, tx1, ty0, ty1, idx, [&](int tag, int ty, int tx) {
            const int im = tag >= n0;
            const int pos = atomicAdd(P.bcur[im] + ty * ntx + tx, 1);
            if (pos < cap) P.blist[im][pos] = tag - (im ? n0 : 0);
        });
    }
}

// Plane coefficients and sorted-corner edge lines of triangle ti of image img.
__device__ __forceinline__ void elas_plane_rec(const ElasPair& P, int img, int ti, float* rec) {
    const int32_t* tr = P.tri[img] + 3 * (size_t)ti;
    const float* sp = P.sp[0];
    const float* sq = P.sp[1];
    int tu[3], tv[3], td[3];
    double A[3][3], b[3];
    int ci[3];
    for (int i = 0; i < 3; i++) ci[i] = tr[i];
    for (int i = 0; i < 3; i++) {
        tu[i] = (int)sp[2 * ci[i]];
        tv[i] = (int)sp[2 * ci[i] + 1];
        td[i] = tu[i] - (int)sq[2 * ci[i]];
        A[i][0] = img ? (double)(tu[i] - td[i]) : (double)tu[i];
        A[i][1] = (double)tv[i];
        A[i][2] = 1.0;
        b[i] = (double)td[i];
    }
    float pa = 0.f, pb = 0.f, pc = 0.f;
    if (elas_solve3(A, b)) { pa = (float)b[0]; pb = (float)b[1]; pc = (float)b[2]; }
    if (img) for (int i = 0; i < 3; i++) tu[i] -= td[i];
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < j; k++)
            if (tu[k] > tu[j]) {
                int x = tu[j]; tu[j] = tu[k]; tu[k] = x;
                x = tv[j]; tv[j] = tv[k]; tv[k] = x;
            }
    const float Au = (float)tu[0], Av = (float)tv[0];
    const float Bu = (float)tu[1], Bv = (float)tv[1];
    const float Cu = (float)tu[2], Cv = (float)tv[2];
    float ABa = 0.f, ACa = 0.f, BCa = 0.f;
    if (tu[0] != tu[1]) ABa = __fdiv_rn(__fsub_rn(Av, Bv), __fsub_rn(Au, Bu));
    if (tu[0] != tu[2]) ACa = __fdiv_rn(__fsub_rn(Av, Cv), __fsub_rn(Au, Cu));
    if (tu[1] != tu[2]) BCa = __fdiv_rn(__fsub_rn(Bv, Cv), __fsub_rn(Bu, Cu));
    rec[0] = pa; rec[1] = pb; rec[2] = pc;
    rec[3] = Au; rec[4] = Av; rec[5] = Bu; rec[6] = Bv; rec[7] = Cu; rec[8] = Cv;
    rec[9] = ABa;  rec[10] = __fsub_rn(Av, __fmul_rn(ABa, Au));
    rec[11] = ACa; rec[12] = __fsub_rn(Av, __fmul_rn(ACa, Au));
    rec[13] = BCa; rec[14] = __fsub_rn(Bv, __fmul_rn(BCa, Bu));
    rec[15] = 0.f;
}

// ---- per-triangle record: KerComputeDisparity.cu:10-124 -------------------
// One lane per triangle of either mesh, 64-lane workgroups so that a few
// thousand triangles spread over many CUs.  Corners come from the left support
// points sp0 for both meshes; the right mesh shifts them by -d (Elas.cu:105,
// KerComputeDisparity.cu:63-66,95-99).  A corner's d is sp0.u - sp1.u (exact
// in float), which saves a dependent load of the lattice.
__global__ __launch_bounds__(64) void k_plane(ElasLaunch L) {
    __shared__ int s_items[5 * 64];
    // the triangle count is data: read the total from the table (a captured
    // launch keeps the grid of its shape bucket)
    const int total = L.prefix[L.npairs];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int idx = (t - L.prefix[p]) * 64 + threadIdx.x;
        const int n0 = P.ntri[0], n1 = P.ntri[1];
        const bool live = idx < n0 + n1;
        const int img = live && idx >= n0;
        const int ti = img ? idx - n0 : idx;
        float* rec = P.trirec[img] + (size_t)ti * ELAS_TRIREC;
        if (live) elas_plane_rec(P, img, ti, rec);
        if (P.gpu_bins) {
            // count the triangle in every tile of its box (k_bin_scan and
            // k_bin_fill follow); the whole wave takes part
            int tx0 = 0, tx1 = 0, ty0 = 0, ty1 = 0;
            const bool ok = live && elas_bin_box(rec, P.W, P.H, tx0, tx1, ty0, ty1);
            const int ntx = P.ntx;
            elas_bin_items(s_items, ok, tx0, tx1, ty0, ty1, idx, [&](int tag, int ty, int tx) {
                atomicAdd(P.bcnt[tag >= n0] + ty * ntx + tx, 1);
            });
        }
    }
}

__device__ __forceinline__ int elas_plane_d(const float* rec, int u, int v) {
    return elas_f2i_sat(__fadd_rn(__fadd_rn(__fmul_rn(rec[0], (float)u), __fmul_rn(rec[1], (float)v)), rec[2]));
}

// ---- rasterisation: KerComputeDisparity.cu:126-150 ------------------------
// One workgroup per 32x32 pixel tile.  The tile's triangles (binned on the
// GPU, or on the host as a fallback) are staged in LDS in chunks; every (triangle, column)
// pair is one work item: it computes the column's half-open row span between
// the long edge AC and AB (u in [A_u, B_u)) or BC (u in [B_u, C_u)) and
// records the triangle id for the span's pixels in an LDS tile with
// atomicMax.  Where spans overlap (the reference races there) the highest
// triangle index wins -- the result of painting the triangles in index order.
// Each pixel's owner then stores the plane prior (int)(a*u + b*v + c) of its
// triangle, or -1.
#define RS_CHUNK 128
__global__ __launch_bounds__(256) void k_raster(ElasLaunch L) {
    __shared__ float s_rec[RS_CHUNK][ELAS_TRIREC];
    __shared__ int s_id[RS_CHUNK];
    __shared__ int s_tile[ELAS_RASTER_T][ELAS_RASTER_T];
    const int tid = threadIdx.x;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int W = P.W, H = P.H;
        const int ntiles = P.ntx * P.nty;
        int local = t - L.prefix[p];
        const int img = local >= ntiles;
        if (img) local -= ntiles;
        const int tx0 = (local % P.ntx) * ELAS_RASTER_T, ty0 = (local / P.ntx) * ELAS_RASTER_T;
        const float* recs = P.trirec[img];
        const int* lst = P.tlist[img];
        const int k0 = elas_ld(P.toff[img] + local), k1 = elas_ld(P.toff[img] + local + 1);
        #pragma unroll
        for (int r = 0; r < 4; r++) s_tile[(tid >> 5) + 8 * r][tid & 31] = -1;
        elas_lds_barrier();
        for (int kc = k0; kc < k1; kc += RS_CHUNK) {
            const int n = min(RS_CHUNK, k1 - kc);
            if (tid < n) {
                const int ti = elas_ld(lst + kc + tid);
                const uint4* src = reinterpret_cast<const uint4*>(recs + (size_t)ti * ELAS_TRIREC);
                const uint4 r0 = elas_ld(src), r1 = elas_ld(src + 1), r2 = elas_ld(src + 2), r3 = elas_ld(src + 3);
                uint4* dst = reinterpret_cast<uint4*>(s_rec[tid]);
                dst[0] = r0; dst[1] = r1; dst[2] = r2; dst[3] = r3;
                s_id[tid] = ti;
            }
            elas_lds_barrier();
            for (int item = tid; item < n * ELAS_RASTER_T; item += 256) {
                const int j = item >> 5, c = item & 31;
                const int u = tx0 + c;
                const float* rec = s_rec[j];
                const int Au = (int)rec[3], Bu = (int)rec[5], Cu = (int)rec[7];
                if (u < Au || u >= Cu || u >= W) continue;
                const bool first = u < Bu;
                const float ea = first ? rec[9] : rec[13], eb = first ? rec[10] : rec[14];
                const float fu = (float)u;
                const int v1 = (int)elas_f2u_sat(__fadd_rn(__fmul_rn(rec[11], fu), rec[12]));
                const int v2 = (int)elas_f2u_sat(__fadd_rn(__fmul_rn(ea, fu), eb));
                const int vlo = max(max(min(v1, v2), 0), ty0);
                const int vhi = min(min(max(v1, v2), H), ty0 + ELAS_RASTER_T);
                const int ti = s_id[j];
                for (int v = vlo; v < vhi; v++) atomicMax(&s_tile[v - ty0][c], ti);
            }
            elas_lds_barrier();
        }
        const int c = tid & 31;
        const int u = tx0 + c;
        if (u < W) {
            #pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = (tid >> 5) + 8 * r, v = ty0 + rr;
                if (v < H) {
                    const int id = s_tile[rr][c];
                    elas_st(P.prior[img] + (size_t)v * W + u,
                            id >= 0 ? elas_plane_d(recs + (size_t)id * ELAS_TRIREC, u, v) : -1);
                }
            }
        }
        elas_lds_barrier();
    }
}

// Debug: the reference's mean map, (float)(int)plane(u,v) where covered, -1.
__global__ __launch_bounds__(256) void k_mean(ElasLaunch L) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int i = (t - L.prefix[p]) * 256 + threadIdx.x;
        if (i >= P.W * P.H) continue;
        P.mean[0][i] = (float)P.prior[0][i];
        P.mean[1][i] = (float)P.prior[1][i];
    }
}

// ---- dense matching: KerComputeDisparity.cu:156-316,378-402 ---------------
// Tile = R image rows (R = min(64/gs, gs), inside one grid-cell row) x
// ELAS_DENSE_WAVES*cpw grid cells; wave w owns cells gx0 + w + WAVES*j, lane =
// (row, column) of a cell, so the cell's candidate set is wave-uniform and is
// walked on the scalar unit as the set bits of its bitmask.  The other image's
// R rows are staged in LDS for the tile's disparity reach, with a row stride
// of 4 (mod 16) descriptors so that a wave's three rows hit disjoint banks;
// the pixel's own 16-byte block stays in registers.  512-thread workgroups
// capped at 64 VGPRs keep two workgroups per CU, so one stages while the
// other matches.
//
// Reference order and tie rule: grid candidates outside the plane window
// [dp-r, dp+r] ascending, then the window ascending with the prior P added,
// strict '<'.  This equals the minimum of the key
//   ((cost + bias) << 11) | (phase << 10) | d,   phase 0 = grid, 1 = window.
// When every P(k) < 0, a grid candidate inside the window is dominated by its
// own window entry (same d, lower cost), so phase 0 need not exclude the
// window; otherwise it does.
template <bool PNEG>
__global__ __launch_bounds__(ELAS_DENSE_THREADS, 8) void k_dense(ElasLaunch L, ElasKParams kp) {
    extern __shared__ uint4 s_i2[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int cpw = kp.dense_cpw;
    const int tcells = ELAS_DENSE_WAVES * cpw;
    // a candidate outside the image (or of an excluded kind) starts its SAD
    // from KEY_BAD: its key exceeds every valid key and stays below 2^32
    constexpr uint32_t KEY_BAD = 0x1ff000u;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int W = P.W, H = P.H, gs = P.gsize, gw = P.gw, dmax = P.dmax;
        const int R = max(1, min(64 / gs, gs));
        const int sub = (gs + R - 1) / R;
        const int segs = (gw + tcells - 1) / tcells;
        const int per_img = P.gh * sub * segs;
        int r = t - L.prefix[p];
        const int img = r / per_img;
        r -= img * per_img;
        const int gy = r / (sub * segs);
        r -= gy * sub * segs;
        const int k = r / segs, s = r - k * segs;
        const int v0 = gy * gs + k * R;
        const int rows = max(0, min(min(R, gs - k * R), H - v0));
        const int gx0 = s * tcells;
        // staged columns: the whole reach of the tile's lanes, [u_lo - dmax,
        // u_lo + tcells*gs) or [u_lo, u_lo + tcells*gs + dmax); columns
        // outside the image are zero and never selected by a valid candidate
        const int u_lo = gx0 * gs;
        const int ncols = tcells * gs + dmax;
        const int c_lo = img ? u_lo : u_lo - dmax;
        const int rs = ncols + ((20 - (ncols & 15)) & 15);  // 4 (mod 16): disjoint banks per row
        const uint8_t* I1 = P.desc[img];
        const uint8_t* I2 = P.desc[1 - img];
        {
            // every thread issues all of its loads before storing any
            const int n = rows * ncols;
            for (int base = 0; base < n; base += 4 * ELAS_DENSE_THREADS) {
                uint4 q[4];
                int li[4];
                int rr = (base + (int)threadIdx.x) / ncols, cc = base + threadIdx.x - rr * ncols;
                #pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int i = base + threadIdx.x + k * ELAS_DENSE_THREADS;
                    const int c = c_lo + cc;
                    q[k] = make_uint4(0u, 0u, 0u, 0u);
                    if (i < n && c >= 0 && c < W)
                        q[k] = elas_ld(reinterpret_cast<const uint4*>(I2 + ((size_t)(v0 + rr) * W + c) * 16));
                    li[k] = i < n ? rr * rs + cc : -1;
                    cc += ELAS_DENSE_THREADS;
                    while (cc >= ncols) {
                        cc -= ncols;
                        rr++;
                    }
                }
                #pragma unroll
                for (int k = 0; k < 4; k++)
                    if (li[k] >= 0) s_i2[li[k]] = q[k];
            }
        }
        // this lane's pixel in each of its cells: prior and own descriptor; lane
        // 32*ci + w holds word w of cell ci's candidate mask
        const int row = lane / gs, col = lane - row * gs;
        uint32_t maskv = 0u;
        {
            const int ci = lane >> 5, w = lane & 31;
            const int gx = gx0 + wave + ci * ELAS_DENSE_WAVES;
            if (ci < cpw && w < P.words && gx < gw) maskv = elas_ld(P.gmask[img] + (size_t)(gy * gw + gx) * P.words + w);
        }
        int dpv[2] = {0, 0};
        uint4 b1v[2] = {make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u)};
        #pragma unroll
        for (int ci = 0; ci < 2; ci++) {
            const int gx = gx0 + wave + ci * ELAS_DENSE_WAVES;
            const int u = gx * gs + col, v = v0 + row;
            const bool valid = ci < cpw && gx < gw && lane < R * gs && row < rows && u < W &&
                               u >= 4 && u <= W - 4 && v >= 4 && v <= H - 4;
            if (valid) {
                dpv[ci] = elas_ld(P.prior[img] + (size_t)v * W + u);
                b1v[ci] = elas_ld(reinterpret_cast<const uint4*>(I1 + ((size_t)v * W + u) * 16));
            }
        }
        elas_lds_barrier();
        #pragma unroll
        for (int ci = 0; ci < 2; ci++) {
            const int gx = gx0 + wave + ci * ELAS_DENSE_WAVES;
            if (ci >= cpw || gx >= gw) continue;
            const int u = gx * gs + col, v = v0 + row;
            const bool active = lane < R * gs && row < rows && u < W;
            const int dp = dpv[ci];
            const uint4 b1 = b1v[ci];
            bool valid = active && u >= 4 && u <= W - 4 && v >= 4 && v <= H - 4 && dp > 0;
            if (kp.match_texture > 0) {
                const uint32_t tex = elas_sad16(b1, make_uint4(0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u), 0u);
                if ((int)tex < kp.match_texture) valid = false;
            }
            // u_warp = u -/+ d must lie in [3, W-4] (KerComputeDisparity.cu:259,267): d <= dlim
            const int dlim = img ? W - 4 - u : u - 3;
            const int sgn = img ? 1 : -1;
            // this lane's descriptor at disparity 0 in the staged row
            const uint4* lp = s_i2 + (row < rows ? row : 0) * rs + (u - c_lo);
            const uint32_t bias = (uint32_t)kp.key_bias;
            const int dpmin = max(dp - kp.plane_radius, 0), dpmax = min(dp + kp.plane_radius, dmax);
            uint32_t kmin = 0xffffffffu;
            // grid candidates (wave-uniform d): key = (sad + bias) << 11 | d
            auto grid_acc = [&](int d) {
                bool ok = d <= dlim;
                if (!PNEG) ok = ok && (d < dpmin || d > dpmax);
                return ok ? bias : KEY_BAD;
            };
            for (int w = 0; w < P.words; w++) {
                uint32_t bits = __builtin_amdgcn_readlane(maskv, 32 * ci + w);
                // four candidates at a time while four remain, so their LDS
                // reads overlap; then one at a time
                while (__builtin_popcount(bits) >= 4) {
                    int dq[4];
                    uint4 b2[4];
                    #pragma unroll
                    for (int q = 0; q < 4; q++) {
                        dq[q] = w * 32 + __builtin_ctz(bits);
                        bits &= bits - 1u;
                        b2[q] = lp[sgn * dq[q]];
                    }
                    #pragma unroll
                    for (int q = 0; q < 4; q++)
                        kmin = min(kmin, (elas_sad16(b1, b2[q], grid_acc(dq[q])) << 11) | (uint32_t)dq[q]);
                }
                while (bits) {
                    const int d = w * 32 + __builtin_ctz(bits);
                    bits &= bits - 1u;
                    kmin = min(kmin, (elas_sad16(b1, lp[sgn * d], grid_acc(d)) << 11) | (uint32_t)d);
                }
            }
            // plane window d = dp - r + k, k = 0..2r, with the prior P(|k - r|)
            // added: key = (sad + P + bias) << 11 | 1024 | d
            const int pr = kp.plane_radius, nwin = 2 * pr + 1;
            const int wlim = min(dmax, dlim);
            const int d0 = dp - pr;
            auto win_key = [&](int k, const uint4 b2) {
                const int d = d0 + k;
                const uint32_t acc = (uint32_t)d <= (uint32_t)wlim ? (uint32_t)((int)bias + kp.P[abs(k - pr)]) : KEY_BAD;
                return (elas_sad16(b1, b2, acc) << 11) | (1024u | ((uint32_t)d & 1023u));
            };
            auto win_lds = [&](int k) { return lp[sgn * min(max(d0 + k, 0), dmax)]; };
            int k0 = 0;
            for (; k0 + 4 <= nwin; k0 += 4) {
                uint4 b2[4];
                #pragma unroll
                for (int q = 0; q < 4; q++) b2[q] = win_lds(k0 + q);
                #pragma unroll
                for (int q = 0; q < 4; q++) kmin = min(kmin, win_key(k0 + q, b2[q]));
            }
            for (; k0 < nwin; k0++) kmin = min(kmin, win_key(k0, win_lds(k0)));
            if (active)
                elas_st(P.draw[img] + (size_t)v * W + u,
                        (valid && kmin < (KEY_BAD << 11)) ? (float)(kmin & 1023u) : -1.0f);
        }
        elas_lds_barrier();
    }
}

// ---- left/right check: KerLeftRightConsistencyCheck.cu:8-66 ---------------
// Reads the dense results (the reference's copies, Elas.cu:541-558) and
// writes the final maps.  |D_other(u -/+ d) - d| > lr_threshold -> -1; an
// invalid partner (-1) passes for d in {0, 1}, as in the reference.
__global__ __launch_bounds__(256) void k_lr(ElasLaunch L, ElasKParams kp) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int i = (t - L.prefix[p]) * 256 + threadIdx.x;
        const int W = P.W;
        if (i >= W * P.H) continue;
        const int v = i / W, u = i - v * W;
        const float* c1 = P.draw[0];
        const float* c2 = P.draw[1];
        const float thr = (float)kp.lr_threshold;
        const float d1 = c1[i], d2 = c2[i];
        const float uw1 = __fsub_rn((float)u, d1), uw2 = __fadd_rn((float)u, d2);
        float o1 = d1, o2 = d2;
        if (d1 >= 0.f && uw1 >= 0.f && uw1 < (float)W) {
            if (fabsf(__fsub_rn(c2[(size_t)v * W + (int)uw1], d1)) > thr) o1 = -1.f;
        } else {
            o1 = -1.f;
        }
        if (d2 >= 0.f && uw2 >= 0.f && uw2 < (float)W) {
            if (fabsf(__fsub_rn(c1[(size_t)v * W + (int)uw2], d2)) > thr) o2 = -1.f;
        } else {
            o2 = -1.f;
        }
        P.dout[0][i] = o1;
        P.dout[1][i] = o2;
    }
}

void elas_launch_bin_clear(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_bin_clear, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_bin_scan(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_bin_scan, dim3(grid), dim3(1024), 0, s, L);
}
void elas_launch_bin_fill(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_bin_fill, dim3(grid), dim3(64), 0, s, L);
}
void elas_launch_plane(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_plane, dim3(grid), dim3(64), 0, s, L);
}
void elas_launch_raster(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_raster, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_mean(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_mean, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_dense(const ElasLaunch& L, const ElasKParams& kp, int grid, size_t lds, hipStream_t s) {
    if (kp.all_p_negative) hipLaunchKernelGGL(k_dense<true>, dim3(grid), dim3(ELAS_DENSE_THREADS), lds, s, L, kp);
    else hipLaunchKernelGGL(k_dense<false>, dim3(grid), dim3(ELAS_DENSE_THREADS), lds, s, L, kp);
}
void elas_launch_lr(const ElasLaunch& L, const ElasKParams& kp, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_lr, dim3(grid), dim3(256), 0, s, L, kp);
}
