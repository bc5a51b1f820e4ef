// k_support.hip -- support-point matching, its filters, compaction and the
// disparity grid (gfx950).
#include <hip/hip_runtime.h>
#include "elas_common.h"
#include "elas_device.h"

//
// Support matching restates src/Kernels/KerSupportMatches.cu:13-192.
// E(d) is the 64-byte SAD of four 16-byte blocks at (+-2, +-2) around the
// candidate (16 x v_sad_u8).
//
// The reference reduces the (best, second best) of 32 threads in thread 0,
// which starts from its own registers and then re-reads its own entry from
// shared memory (:154-178), so thread 0's best is counted twice.  Thread t
// holds d with (d - dmin) % 32 == t.  The resulting decision, in closed form:
//   m      = min E,   (r*, d*) = lexicographic min of ((d - dmin) & 31, d)
//            over the minimisers (the first entry of the serial scan);
//   min2   = m if r* == 0 (thread 0 holds a minimiser: duplicated), else the
//            second smallest E counted with multiplicity;
//   accept = (float)m < support_threshold * (float)min2,   winner d*.
// With key(d) = E << 16 | r << 10 | d (distinct per d), m, r*, d* come from
// the smallest key and the second smallest E from the second smallest key.
//
// Mapping: one workgroup (4 waves) = 64 consecutive candidates of a lattice
// row.  Lane (j, c) of wave w works on candidate 16w + c (in the right pass
// the candidate dealt to that place, see k_support_lr) and the disparities
// d = dmin + j + 4m (j = lane / 16).  The other image's descriptor rows v-2
// and v+2 are staged in LDS over the columns the 64 candidates can reach.
// E(d) pairs the reference blocks with the block pair X(d) at column
// u -/+ d -/+ 2 and with X(d -/+ 4) at the column 4 further in: stepping d by
// 4, each lane loads one new block pair and reuses the previous one, so LDS
// supplies 32 bytes per (candidate, d).  Within a 16-lane group the
// candidates' columns are 16*step bytes apart (no bank conflicts for step 5).
// The four lanes of a candidate merge their two smallest keys with two
// shuffles.
template <bool RIGHT>
__device__ __forceinline__ int support_pass(const ElasPair& P, const ElasKParams& kp, uint4* s_rows, int v_can,
                                            int k0, int cidx, int d_left_in) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int j = lane >> 4;
    {
        const int W = P.W, H = P.H, cw = P.cw, step = P.step;
        const int u_can = k0 + cidx;  // cidx < 0: no candidate
        const int v = v_can * step;
        const int u_first = k0 * step, u_last = min(k0 + 63, cw - 1) * step;
        const int c_lo = max(u_first - P.dmax - 2, 0);
        const int c_hi = RIGHT ? min(u_last + P.dmax + 3, W) : min(u_last + 3, W);
        const int ncols = max(c_hi - c_lo, 0);
        const bool row_ok = v >= 5 && v <= H - 6;
        const uint8_t* I1 = P.desc[RIGHT ? 1 : 0] + (size_t)16 * W * v;
        if (row_ok) {
            // every thread issues all of its loads (buffer loads: a lane past
            // the row reads nothing, without a branch) before storing any;
            // rows v-2 and v+2 side by side
            const __amdgpu_buffer_rsrc_t rr = elas_rsrc((void*)P.desc[RIGHT ? 0 : 1], 16u * (uint32_t)W * (uint32_t)H);
            const uint32_t top0 = 16u * ((uint32_t)(v - 2) * (uint32_t)W + (uint32_t)c_lo);
            const uint32_t bot0 = top0 + 64u * (uint32_t)W;
            for (int base = 0; base < ncols; base += 4 * 256) {
                uint4 q[8];
                #pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int i = base + tid + k * 256;
                    q[k] = elas_ld16_ranged(rr, i < ncols ? top0 + 16u * (uint32_t)i : ELAS_OOR);
                    q[4 + k] = elas_ld16_ranged(rr, i < ncols ? bot0 + 16u * (uint32_t)i : ELAS_OOR);
                }
                #pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int i = base + tid + k * 256;
                    if (i < ncols) {
                        s_rows[i] = q[k];
                        s_rows[ncols + i] = q[4 + k];
                    }
                }
            }
        }
        int u = u_can * step;
        bool ok = row_ok && cidx >= 0 && u_can < cw && u >= 5 && u <= W - 6;
        const int d_left = RIGHT ? d_left_in : 0;
        if (ok && RIGHT) {
            if (d_left < 0) ok = false;
            else u -= d_left;
        }
        const int dminv = max(P.dmin, 0);
        const int dmaxv = RIGHT ? min(P.dmax, W - u - 5) : min(P.dmax, u - 5);
        uint4 rTL, rTR, rBL, rBR;
        if (ok) {
            const uint8_t* b1 = I1 + 16 * (size_t)u;
            const uint4 c = elas_ld(reinterpret_cast<const uint4*>(b1));
            const uint32_t tex = elas_sad16(c, make_uint4(0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u), 0u);
            ok = (int)tex >= kp.support_texture && dmaxv - dminv >= 10;
            const long rowoff = 32L * W;
            rTL = elas_ld(reinterpret_cast<const uint4*>(b1 - 32 - rowoff));
            rTR = elas_ld(reinterpret_cast<const uint4*>(b1 + 32 - rowoff));
            rBL = elas_ld(reinterpret_cast<const uint4*>(b1 - 32 + rowoff));
            rBR = elas_ld(reinterpret_cast<const uint4*>(b1 + 32 + rowoff));
        }
        elas_lds_barrier();
        uint32_t k1 = 0xffffffffu, k2 = 0xffffffffu;
        if (ok) {
            // block pair at column cx(d) = u - d - 2 (left) / u + d + 2 (right),
            // and the pair at cx(d) + 4 / cx(d) - 4, which is cx(d - 4)
            const uint4 rXt = RIGHT ? rTR : rTL, rXb = RIGHT ? rBR : rBL;
            const uint4 rPt = RIGHT ? rTL : rTR, rPb = RIGHT ? rBL : rBR;
            const uint4* top = s_rows - c_lo;
            const uint4* bot = s_rows + ncols - c_lo;
            int d = dminv + j;
            int cx = RIGHT ? u + d + 2 : u - d - 2;
            const int dstep = RIGHT ? 4 : -4;
            // (k1, k2) <- the two smallest of (k1, k2, key), k1 <= k2: the new
            // k2 is the median of the three
            auto take = [&](uint32_t key) {
                k2 = max(min(k1, k2), min(max(k1, k2), key));
                k1 = min(k1, key);
            };
            // A lane scans d = dmin + j + 4 m, m = 0, 1, ..  Within the scan its
            // key is E << 16 | low(m) with low(m) = (m mod 8) << 7 | m / 8: the
            // same order as the final key's (r, d) with r = (d - dmin) mod 32
            // = j + 4 (m mod 8), and uniform across the lanes, so it enters
            // the SAD chain (v_sad_hi_u8 adds each block's SAD << 16) as the
            // scalar start value.
            auto low = [](int m) { return (uint32_t)(((m & 7) << 7) | (m >> 3)); };
            auto energy = [&](const uint4 xt, const uint4 xb, const uint4 pt, const uint4 pb, uint32_t c) {
                c = elas_sad16_hi(rXt, xt, c);
                c = elas_sad16_hi(rXb, xb, c);
                c = elas_sad16_hi(rPt, pt, c);
                return elas_sad16_hi(rPb, pb, c);
            };
            const int n = d <= dmaxv ? (dmaxv - d) / 4 + 1 : 0;
            const int n0 = __builtin_amdgcn_readfirstlane(n);
            // the pair of the previous step, carried
            uint4 pt = top[cx - dstep], pb = bot[cx - dstep];
            if (__ballot(n != n0) == 0) {
                // the usual case: every computing lane of the wave walks the same
                // count, so the loop runs on the scalar unit
                int i = 0;
                // eight steps an iteration, as two groups of four whose loads
                // are not hoisted above the other group's SADs: the second
                // group's last pair can take the registers of the carried
                // pair, so the carry costs no copies
                for (; i + 8 <= n0; i += 8) {
                    uint4 xt[4], xb[4];
                    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        xt[k] = top[cx + k * dstep];
                        xb[k] = bot[cx + k * dstep];
                    }
                    #pragma unroll
                    for (int k = 0; k < 4; k++)
                        take(energy(xt[k], xb[k], k ? xt[k - 1] : pt, k ? xb[k - 1] : pb, low(i + k)));
                    __builtin_amdgcn_sched_barrier(0);
                    uint4 yt[4], yb[4];
                    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        yt[k] = top[cx + (4 + k) * dstep];
                        yb[k] = bot[cx + (4 + k) * dstep];
                    }
                    #pragma unroll
                    for (int k = 0; k < 4; k++)
                        take(energy(yt[k], yb[k], k ? yt[k - 1] : xt[3], k ? yb[k - 1] : xb[3], low(i + 4 + k)));
                    __builtin_amdgcn_sched_barrier(0);
                    pt = yt[3];
                    pb = yb[3];
                    cx += 8 * dstep;
                }
                for (; i + 4 <= n0; i += 4) {
                    uint4 xt[4], xb[4];
                    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        xt[k] = top[cx + k * dstep];
                        xb[k] = bot[cx + k * dstep];
                    }
                    #pragma unroll
                    for (int k = 0; k < 4; k++)
                        take(energy(xt[k], xb[k], k ? xt[k - 1] : pt, k ? xb[k - 1] : pb, low(i + k)));
                    pt = xt[3];
                    pb = xb[3];
                    cx += 4 * dstep;
                }
                for (; i < n0; i++, cx += dstep) {
                    const uint4 xt = top[cx], xb = bot[cx];
                    take(energy(xt, xb, pt, pb, low(i)));
                    pt = xt;
                    pb = xb;
                }
            } else {
                // lanes walk different counts (near the image border): the
                // largest count in groups of eight steps, a lane's steps past
                // its own count starting from 0xC0000000, which stays above
                // every real key (energies < 2^14) and below 0xffffffff, and
                // never ranks among a candidate's two smallest (every
                // computing candidate has at least 11 disparities)
                int nmax = n0;
                for (uint64_t more = __ballot(n > nmax); more; more = __ballot(n > nmax))
                    nmax = __builtin_amdgcn_readlane(n, __builtin_ctzll(more));
                for (int i = 0; i < nmax; i += 8) {
                    uint4 xt[4], xb[4];
                    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        xt[k] = top[cx + k * dstep];
                        xb[k] = bot[cx + k * dstep];
                    }
                    #pragma unroll
                    for (int k = 0; k < 4; k++)
                        take(energy(xt[k], xb[k], k ? xt[k - 1] : pt, k ? xb[k - 1] : pb,
                                    i + k < n ? low(i + k) : 0xC0000000u));
                    __builtin_amdgcn_sched_barrier(0);
                    uint4 yt[4], yb[4];
                    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        yt[k] = top[cx + (4 + k) * dstep];
                        yb[k] = bot[cx + (4 + k) * dstep];
                    }
                    #pragma unroll
                    for (int k = 0; k < 4; k++)
                        take(energy(yt[k], yb[k], k ? yt[k - 1] : xt[3], k ? yb[k - 1] : xb[3],
                                    i + 4 + k < n ? low(i + 4 + k) : 0xC0000000u));
                    __builtin_amdgcn_sched_barrier(0);
                    pt = yt[3];
                    pb = yb[3];
                    cx += 8 * dstep;
                }
            }
            // final keys E << 16 | r << 10 | d, r = (d - dmin) mod 32 (r < 32 and
            // d < 1024 fill disjoint bits)
            auto final_key = [&](uint32_t k) {
                const uint32_t lo = k & 0xffffu;
                const int m = (int)(((lo & 127u) << 3) | (lo >> 7));
                const uint32_t r = (uint32_t)(j + 4 * (m & 7)), dd = (uint32_t)(dminv + j + 4 * m);
                return k == 0xffffffffu ? k : (k & 0xffff0000u) | (r << 10) | dd;
            };
            k1 = final_key(k1);
            k2 = final_key(k2);
        }
        #pragma unroll
        for (int off = 16; off <= 32; off <<= 1) {
            const uint32_t o1 = __shfl_xor(k1, off, 64);
            const uint32_t o2 = __shfl_xor(k2, off, 64);
            k2 = min(max(k1, o1), min(k2, o2));
            k1 = min(k1, o1);
        }
        int result = -1;
        if (ok) {
            const uint32_t m = k1 >> 16;
            const uint32_t min2 = (((k1 >> 10) & 31u) == 0u) ? m : (k2 >> 16);
            const int dstar = (int)(k1 & 1023u);
            if ((float)m < kp.support_threshold * (float)min2) {
                if (!RIGHT) result = dstar;
                else result = (abs(d_left - dstar) > kp.lr_threshold) ? -1 : d_left;
            }
        }
        elas_lds_barrier();  // s_rows is free for the next pass
        return result;
    }
}

// Both passes of a unit in one workgroup.  The left pass's result of a
// candidate stays in LDS and feeds the right-to-left check, which stages the
// other image's rows in the same LDS.  The right pass walks the columns
// u - d_left + d + 2, so neighbouring candidates whose d_left differ start
// their walks in colliding LDS bank groups (a 16-byte column is bank group
// column mod 16, and the walk moves every lane of a wave by the same amount).
// Its candidates (those with a left match) are therefore dealt to the waves
// in order of that residue: rank r goes to wave r mod nw, lane r / nw, with
// nw = ceil(count / 16), so a wave's candidates have distinct residues unless
// more than nw share one, and the waves beyond nw have nothing to do.  The
// results do not depend on the placement.
__global__ __launch_bounds__(256) void k_support_lr(ElasLaunch L, ElasKParams kp) {
    extern __shared__ uint4 s_rows[];  // [2][ncols]: row v-2, then row v+2
    __shared__ int32_t s_dl[64];       // the unit's left results
    __shared__ int32_t s_map[64];      // right pass: the candidate at 16 w + c, or -1
    __shared__ int32_t s_bin[17];      // residue counts, then their starts; [16]: candidates without a match
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = 16 * wave + (lane & 15);  // this lane's candidate in the left pass
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int cw = P.cw;
        const int local = t - L.prefix[p];
        const int per_row = (cw + 63) >> 6;
        const int v_can = local / per_row;
        const int k0 = (local - v_can * per_row) * 64;
        const int dl = support_pass<false>(P, kp, s_rows, v_can, k0, c, 0);
        if (tid < 17) s_bin[tid] = 0;
        if (tid < 64) s_map[tid] = -1;
        elas_lds_barrier();
        const bool lead = (lane >> 4) == 0;  // one lane per candidate
        const int key = dl >= 0 ? ((P.step * (k0 + c) - dl) & 15) : 16;
        int rank = 0;
        if (lead) {
            s_dl[c] = dl;
            rank = atomicAdd(&s_bin[key], 1);
        }
        elas_lds_barrier();
        if (tid == 0) {
            int acc = 0;
            for (int k = 0; k < 17; k++) {
                const int n = s_bin[k];
                s_bin[k] = acc;
                acc += n;
            }
        }
        elas_lds_barrier();
        if (lead && key < 16) {
            const int r = s_bin[key] + rank, nw = max(1, (s_bin[16] + 15) >> 4);
            s_map[16 * (r % nw) + r / nw] = c;
        }
        elas_lds_barrier();
        const int cr = s_map[c];
        const int dr = support_pass<true>(P, kp, s_rows, v_can, k0, cr, cr >= 0 ? s_dl[cr] : -1);
        if (lead) {
            if (k0 + c < cw) {
                elas_st(P.can[0] + v_can * cw + k0 + c, dl);
                if (dl < 0) elas_st(P.can[1] + v_can * cw + k0 + c, -1);
            }
            if (cr >= 0) elas_st(P.can[1] + v_can * cw + k0 + cr, dr);
        }
    }
}

// Inconsistency filter, KerSupportMatches.cu:196-238, read from can[1]
// (snapshot) and written to can[2].  The reference updates in place and
// races; reading a snapshot makes the result independent of scheduling.
// The row guard y < can_width is the reference's (:208).  A 16x16-cell tile
// stages its (16+2R)^2 window in LDS (R <= 8; wider windows read L2).
#define INC_T 16
#define INC_RMAX 8
__global__ __launch_bounds__(256) void k_incon(ElasLaunch L, ElasKParams kp) {
    __shared__ int s_win[INC_T + 2 * INC_RMAX][INC_T + 2 * INC_RMAX + 1];
    const int tid = threadIdx.x;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int cw = P.cw, ch = P.ch;
        const int local = t - L.prefix[p];
        const int tx = (cw + INC_T - 1) / INC_T;
        const int x0 = (local % tx) * INC_T, y0 = (local / tx) * INC_T;
        const int R = kp.incon_window;
        const bool lds = R <= INC_RMAX;
        const int32_t* in = P.can[1];
        if (lds) {
            // all loads of a thread in flight before its stores; rows of the
            // window by incremental index (no division)
            const int side = INC_T + 2 * R;
            constexpr int K = ((INC_T + 2 * INC_RMAX) * (INC_T + 2 * INC_RMAX) + 255) / 256;
            int q[K];
            int yy = tid / side, xx = tid - yy * side;
            const int qy = 256 / side, rx = 256 - qy * side;
            #pragma unroll
            for (int k = 0; k < K; k++) {
                const int gx = x0 - R + xx, gy = y0 - R + yy;
                q[k] = (yy < side && gx >= 0 && gx < cw && gy >= 0 && gy < ch) ? elas_ld(in + gy * cw + gx) : -1;
                yy += qy;
                xx += rx;
                if (xx >= side) { xx -= side; yy++; }
            }
            yy = tid / side;
            xx = tid - yy * side;
            #pragma unroll
            for (int k = 0; k < K; k++) {
                if (yy < side) s_win[yy][xx] = q[k];
                yy += qy;
                xx += rx;
                if (xx >= side) { xx -= side; yy++; }
            }
        }
        elas_lds_barrier();
        const int x = x0 + (tid & (INC_T - 1)), y = y0 + (tid / INC_T);
        if (x < cw && y < ch) {
            const int lx = x - x0, ly = y - y0;
            const int d = lds ? s_win[ly + R][lx + R] : in[y * cw + x];
            int res = d;
            if (d >= 0 && y < cw) {
                int support = 0;
                // |d - d2| <= thr  <=>  (unsigned)(d2 - d + thr) <= 2 thr, for thr >= 0
                const unsigned thr2 = 2u * (unsigned)kp.incon_threshold;
                if (lds && R == 5 && kp.incon_threshold >= 0) {
                    #pragma unroll
                    for (int dy = 0; dy <= 10; dy++)
                        #pragma unroll
                        for (int dx = 0; dx <= 10; dx++) {
                            const int d2 = s_win[ly + dy][lx + dx];
                            support += (d2 >= 0 && (unsigned)(d2 - d + kp.incon_threshold) <= thr2) ? 1 : 0;
                        }
                } else if (lds) {
                    for (int dy = 0; dy <= 2 * R; dy++)
                        for (int dx = 0; dx <= 2 * R; dx++) {
                            const int d2 = s_win[ly + dy][lx + dx];
                            support += (d2 >= 0 && abs(d - d2) <= kp.incon_threshold) ? 1 : 0;
                        }
                } else {
                    const int xa = max(x - R, 0), xb = min(x + R, cw - 1);
                    const int ya = max(y - R, 0), yb = min(y + R, ch - 1);
                    for (int y2 = ya; y2 <= yb; y2++)
                        for (int x2 = xa; x2 <= xb; x2++) {
                            const int d2 = in[y2 * cw + x2];
                            support += (d2 >= 0 && abs(d - d2) <= kp.incon_threshold) ? 1 : 0;
                        }
                }
                if (support < kp.incon_min_support) res = -1;
            }
            P.can[2][y * cw + x] = res;
        }
        elas_lds_barrier();
    }
}

// Redundancy filters, KerSupportMatches.cu:242-356: REDUN_MAX_DIST 5,
// REDUN_THRESHOLD 1, horizontal (can[2] -> can[3]) then vertical (in place on
// can[3]).  The reference scans each line in order and removes a point when a
// valid point within 5 cells on BOTH sides differs by at most 1.  In that
// sequential scan the side ahead still holds the input values and the side
// behind holds the results, so for cell x
//   ahead(x)  = some x+j, j in 1..5, is valid and within 1     (input only)
//   M(x)      = bits j-1 for which x-j is valid and within 1   (input only)
//   removed(x)= valid(x) && ahead(x) && (M(x) & ~H(x)) != 0,
// where H(x) holds the removed flags of x-1..x-5.  A workgroup takes lpw lines
// (elas_redun_lines) into LDS with coalesced loads, computes ahead/M for every
// cell in parallel, then one wave per line runs the 5-bit recurrence (16-cell
// segments for all 32 entry states in parallel, chained by a scalar walk), and
// all threads write the lines back.
#define RD_INV (-30000)  // invalid cell in the int16 strip: never within 1 of a valid d
__device__ __forceinline__ bool rd_near(int a, int d) { return (unsigned)(a - d + 1) <= 2u; }

template <bool VERTICAL>
__global__ __launch_bounds__(256) void k_redun(ElasLaunch L) {
    extern __shared__ int16_t s_rd[];  // [lpw][len] values, then [lpw][lenp] flag bytes
    const int tid = threadIdx.x;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int cw = P.cw, ch = P.ch;
        const int nlines = VERTICAL ? cw : ch, len = VERTICAL ? ch : cw;
        const int lpw = elas_redun_lines(len);
        const int lenp = (len + 15) & ~15;
        const int line0 = (t - L.prefix[p]) * lpw;
        const int nl = min(lpw, nlines - line0);
        const int32_t* in = VERTICAL ? P.can[3] : P.can[2];
        int32_t* out = P.can[3];
        int16_t* s_val = s_rd;
        uint8_t* s_flag = reinterpret_cast<uint8_t*>(s_rd) + ((2 * lpw * len + 15) & ~15);
        // cell e of the strip: H -> (line e / len, pos e % len), V -> (pos e / nl, line e % nl)
        const int n = nl * len;
        const int inner = VERTICAL ? nl : len;
        auto gaddr = [&](int a, int b) -> size_t {  // (outer, inner) -> lattice index
            return VERTICAL ? (size_t)a * cw + line0 + b : (size_t)(line0 + a) * cw + b;
        };
        auto laddr = [&](int a, int b) { return VERTICAL ? b * len + a : a * len + b; };
        const int qa = 256 / inner, rb = 256 - qa * inner;  // a step of 256 cells
        for (int base = 0; base < n; base += 16 * 256) {
            int q[16], li[16];
            int a = (base + tid) / inner, b = base + tid - a * inner;
            #pragma unroll
            for (int k = 0; k < 16; k++) {
                const bool ok = base + tid + 256 * k < n;
                q[k] = ok ? elas_ld(in + gaddr(a, b)) : -1;
                li[k] = ok ? laddr(a, b) : -1;
                a += qa;
                b += rb;
                if (b >= inner) {
                    b -= inner;
                    a++;
                }
            }
            #pragma unroll
            for (int k = 0; k < 16; k++)
                if (li[k] >= 0) s_val[li[k]] = (int16_t)(q[k] < 0 ? RD_INV : q[k]);
        }
        elas_lds_barrier();
        // flag byte: M(x) if valid(x) && ahead(x), else 0 (x is then never
        // removed); a lane takes 2 cells (consecutive lanes read consecutive
        // words: no bank conflicts)
        const int runs = lenp >> 1;  // flags cover the padding too (as 0)
        for (int e = tid; e < nl * runs; e += 256) {
            const int l = e / runs, x0 = (e - l * runs) * 2;
            const int16_t* row = s_val + l * len;
            int w[12];
            #pragma unroll
            for (int j = 0; j < 12; j++) {
                const int x = x0 - 5 + j;
                w[j] = (x >= 0 && x < len) ? row[x] : RD_INV;
            }
            #pragma unroll
            for (int i = 0; i < 2; i++) {
                const int d = w[i + 5];
                uint32_t f = 0;
                bool ahead = false;
                #pragma unroll
                for (int j = 1; j <= 5; j++) {
                    f |= rd_near(w[i + 5 - j], d) ? (1u << (j - 1)) : 0u;
                    ahead |= rd_near(w[i + 5 + j], d);
                }
                s_flag[l * lenp + x0 + i] = (uint8_t)((d >= 0 && ahead) ? f : 0u);
            }
        }
        elas_lds_barrier();
        {
            // The recurrence, one line per wave.  nh has bit j-1 set when x-j
            // was kept; x is removed when its flag meets a kept neighbour, and
            // nh becomes ((nh << 1) & 30) | !removed.  A lane takes a segment of
            // 16 cells and runs them for all 32 entry states at once: plane b
            // holds bit b of nh for entry state s in its bit s, and the
            // removals in bit s of one word per cell.  A scalar walk over the
            // segments (five readlanes each) then finds every segment's actual
            // entry state, and each lane applies its removals.
            const int wv = tid >> 6, lane = tid & 63;
            if (wv < nl) {
                const uint8_t* fl = s_flag + wv * lenp;
                int16_t* row = s_val + wv * len;
                const int nseg = lenp >> 4;
                uint32_t s_in = 0;  // the state entering the first segment of the chunk
                for (int g0 = 0; g0 < nseg; g0 += 64) {
                    const int seg = g0 + lane;
                    const uint4 f = seg < nseg ? *reinterpret_cast<const uint4*>(fl + 16 * seg) : make_uint4(0, 0, 0, 0);
                    const uint32_t fw[4] = {f.x, f.y, f.z, f.w};
                    uint32_t p0 = 0xAAAAAAAAu, p1 = 0xCCCCCCCCu, p2 = 0xF0F0F0F0u, p3 = 0xFF00FF00u, p4 = 0xFFFF0000u;
                    uint32_t rp[16];
                    #pragma unroll
                    for (int c = 0; c < 16; c++) {
                        const uint32_t m = (fw[c >> 2] >> (8 * (c & 3))) & 31u;
                        const uint32_t r = ((m & 1u) ? p0 : 0u) | ((m & 2u) ? p1 : 0u) | ((m & 4u) ? p2 : 0u) |
                                           ((m & 8u) ? p3 : 0u) | ((m & 16u) ? p4 : 0u);
                        rp[c] = r;
                        p4 = p3;
                        p3 = p2;
                        p2 = p1;
                        p1 = p0;
                        p0 = ~r;
                    }
                    const int ns = min(64, nseg - g0);
                    uint32_t st = s_in, my_in = 0;
                    for (int k = 0; k < ns; k++) {
                        my_in = lane == k ? st : my_in;
                        const uint32_t q0 = __builtin_amdgcn_readlane(p0, k), q1 = __builtin_amdgcn_readlane(p1, k),
                                       q2 = __builtin_amdgcn_readlane(p2, k), q3 = __builtin_amdgcn_readlane(p3, k),
                                       q4 = __builtin_amdgcn_readlane(p4, k);
                        st = ((q0 >> st) & 1u) | (((q1 >> st) & 1u) << 1) | (((q2 >> st) & 1u) << 2) |
                             (((q3 >> st) & 1u) << 3) | (((q4 >> st) & 1u) << 4);
                    }
                    s_in = st;
                    if (seg < nseg) {
                        #pragma unroll
                        for (int c = 0; c < 16; c++)
                            if (((rp[c] >> my_in) & 1u) && 16 * seg + c < len) row[16 * seg + c] = (int16_t)RD_INV;
                    }
                }
            }
        }
        elas_lds_barrier();
        {
            int a = tid / inner, b = tid - a * inner;
            for (int e = tid; e < n; e += 256) {
                const int v = s_val[laddr(a, b)];
                out[gaddr(a, b)] = v < 0 ? -1 : v;
                a += qa;
                b += rb;
                if (b >= inner) {
                    b -= inner;
                    a++;
                }
            }
        }
        elas_lds_barrier();
    }
}

// Compaction, KerSupportMatches.cu:361-423 + Elas.cu:338-394: row-major
// order.  Pass 1 counts valid points per 256-cell block; pass 2 adds the
// counts of earlier blocks and writes sp0 = (u*step, v*step) and
// sp1 = (u*step - d, v*step); the last block stores the total.  Points and
// total go to the device arena and, for the host Delaunay step, to mapped
// pinned host memory (so no copy sits between phase A and Triangle).
__global__ __launch_bounds__(256) void k_blkcount(ElasLaunch L) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int b = t - L.prefix[p];
        const int idx = b * 256 + threadIdx.x;
        const bool valid = idx < P.cw * P.ch && P.can[3][idx] >= 0;
        const int n = __syncthreads_count(valid);
        if (threadIdx.x == 0) P.blkcnt[b] = n;
    }
}

__global__ __launch_bounds__(256) void k_compact(ElasLaunch L) {
    __shared__ int s_red[256];
    __shared__ int s_wave[4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int b = t - L.prefix[p];
        int acc = 0;
        for (int j = tid; j < b; j += 256) acc += P.blkcnt[j];
        s_red[tid] = acc;
        elas_lds_barrier();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) s_red[tid] += s_red[tid + s];
            elas_lds_barrier();
        }
        const int base = s_red[0];
        const int idx = b * 256 + tid;
        const int d = idx < P.cw * P.ch ? P.can[3][idx] : -1;
        const bool valid = d >= 0;
        const uint64_t mask = __ballot(valid);
        const int before = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) s_wave[wave] = __popcll(mask);
        elas_lds_barrier();
        int off = base + before;
        for (int w = 0; w < wave; w++) off += s_wave[w];
        if (valid) {
            const int vc = idx / P.cw, uc = idx - vc * P.cw;
            const float pu = (float)(uc * P.step), pv = (float)(vc * P.step);
            reinterpret_cast<float2*>(P.sp[0])[off] = make_float2(pu, pv);
            reinterpret_cast<float2*>(P.sp[1])[off] = make_float2(pu - (float)d, pv);
            reinterpret_cast<float2*>(P.hsp[0])[off] = make_float2(pu, pv);
            reinterpret_cast<float2*>(P.hsp[1])[off] = make_float2(pu - (float)d, pv);
        }
        if (b == P.nblk - 1 && tid == 0) {
            const int total = base + s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
            P.nsp[0] = total;
            P.hnsp[0] = total;
        }
        elas_lds_barrier();
    }
}

// Disparity grid, KerGrid.cu:9-126 + Elas.cu:403-464.  Flags are bitmasks
// (words x 32 disparities per cell): fill, scatter d-1..d+1 of each support
// point with atomicOr, then OR the 3x3 neighbourhood of interior cells and
// write its set bits as the cell's candidate list, ascending (the reference's
// list, KerGrid.cu:112-124), in 16-bit entries behind the count and the
// largest d.  k_grid_list renders the reference's int32 layout ([0] = count,
// [1..] = d, zero tail) for parity dumps.
__global__ __launch_bounds__(256) void k_grid_fill(ElasLaunch L) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int n = P.gw * P.gh * P.words;
        const int i = (t - L.prefix[p]) * 256 + threadIdx.x;
        if (i < 2 * n) P.gflag[i >= n][i >= n ? i - n : i] = 0u;
        // the raster's culling lists of phase B start empty
        if (t == L.prefix[p])
            for (int s = threadIdx.x; s < 2 * P.nst; s += 256) P.scnt[s >= P.nst][s >= P.nst ? s - P.nst : s] = 0;
    }
}

__global__ __launch_bounds__(256) void k_grid_scatter(ElasLaunch L) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int idx = (t - L.prefix[p]) * 256 + threadIdx.x;
        if (idx >= P.cw * P.ch) continue;
        const int d = P.can[3][idx];
        if (d < 0) continue;
        const int vc = idx / P.cw, uc = idx - vc * P.cw;
        const int u_img = uc * P.step, v_img = vc * P.step;
        const int ug0 = elas_floordiv(u_img, P.gsize);
        const int ug1 = elas_floordiv(u_img - d, P.gsize);
        const int vg = elas_floordiv(v_img, P.gsize);
        if (vg < 0 || vg >= P.gh) continue;
        const int dlo = max(d - 1, 0), dhi = min(d + 1, P.dmax);
        for (int dd = dlo; dd <= dhi; dd++) {
            const uint32_t bit = 1u << (dd & 31);
            if (ug0 >= 0 && ug0 < P.gw) atomicOr(&P.gflag[0][(vg * P.gw + ug0) * P.words + (dd >> 5)], bit);
            if (ug1 >= 0 && ug1 < P.gw) atomicOr(&P.gflag[1][(vg * P.gw + ug1) * P.words + (dd >> 5)], bit);
        }
    }
}

__global__ __launch_bounds__(256) void k_grid_dilate(ElasLaunch L) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int ncell = P.gw * P.gh;
        const int i = (t - L.prefix[p]) * 256 + threadIdx.x;
        if (i >= 2 * ncell) continue;
        const int img = i >= ncell, c = img ? i - ncell : i;
        const int gy = c / P.gw, gx = c - gy * P.gw;
        const bool interior = gx > 0 && gx < P.gw - 1 && gy > 0 && gy < P.gh - 1;
        const uint32_t* f = P.gflag[img];
        uint16_t* g = P.gcand[img] + (size_t)c * P.cstride;
        int n = 0, last = 0;
        if (interior)
            for (int w = 0; w < P.words; w++) {
                uint32_t acc = 0u;
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++)
                        acc |= f[((gy + dy) * P.gw + (gx + dx)) * P.words + w];
                for (; acc; acc &= acc - 1u) {
                    last = 32 * w + __builtin_ctz(acc);
                    g[2 + n++] = (uint16_t)last;
                }
            }
        g[0] = (uint16_t)n;
        g[1] = (uint16_t)last;
    }
}

__global__ __launch_bounds__(256) void k_grid_list(ElasLaunch L) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int ncell = P.gw * P.gh;
        const int i = (t - L.prefix[p]) * 256 + threadIdx.x;
        if (i >= 2 * ncell) continue;
        const int img = i >= ncell, c = img ? i - ncell : i;
        const uint16_t* m = P.gcand[img] + (size_t)c * P.cstride;
        int32_t* g = P.glist[img] + (size_t)c * (P.dmax + 2);
        const int n = m[0];
        for (int k = 0; k < n; k++) g[1 + k] = m[2 + k];
        for (int k = 1 + n; k < P.dmax + 2; k++) g[k] = 0;
        g[0] = n;
    }
}

void elas_launch_support(const ElasLaunch& L, const ElasKParams& kp, int grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_support_lr, dim3(grid), dim3(256), lds, s, L, kp);
}
void elas_launch_incon(const ElasLaunch& L, const ElasKParams& kp, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_incon, dim3(grid), dim3(256), 0, s, L, kp);
}
void elas_launch_redun(const ElasLaunch& L, bool vertical, int grid, size_t lds, hipStream_t s) {
    if (vertical) hipLaunchKernelGGL(k_redun<true>, dim3(grid), dim3(256), lds, s, L);
    else hipLaunchKernelGGL(k_redun<false>, dim3(grid), dim3(256), lds, s, L);
}
void elas_launch_blkcount(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_blkcount, dim3(grid), dim3(256), 0, s, L);
}
// ---- corner support points (add_corners, src/Elas.h:23) ---------------------
// The reference declares the switch without implementing it; its comment
// ("add support points at image corners with nearest neighbor disparities")
// is the specification.  One workgroup per pair, after compaction: each image
// corner (0,0), (0,H-1), (W-1,0), (W-1,H-1) takes the disparity of the
// support point nearest to it (smallest squared distance, the first in
// compaction order on ties: a minimum of (distance, index) keys), and six
// points are appended after the nsp support points: the four corners, then
// the two right corners again at u = W-1+d (so that the right image's mesh,
// over (u-d, v), reaches its right border).  The count becomes nsp + 6, in
// device memory and in the pinned copy the host step reads.
__global__ __launch_bounds__(256) void k_corners(ElasLaunch L) {
    __shared__ unsigned long long red[4][256];
    const int tid = threadIdx.x;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int n = P.nsp[0];
        if (n < 1) continue;  // uniform: no support point, no corners
        const int cu[4] = {0, 0, P.W - 1, P.W - 1}, cv[4] = {0, P.H - 1, 0, P.H - 1};
        unsigned long long key[4] = {~0ull, ~0ull, ~0ull, ~0ull};
        const float2* s0 = reinterpret_cast<const float2*>(P.sp[0]);
        for (int i = tid; i < n; i += 256) {
            const float2 q = s0[i];
            const int u = (int)q.x, v = (int)q.y;
            for (int c = 0; c < 4; c++) {
                const unsigned du = (unsigned)abs(u - cu[c]), dv = (unsigned)abs(v - cv[c]);
                const unsigned long long k2 = ((unsigned long long)(du * du + dv * dv) << 32) | (unsigned)i;
                key[c] = k2 < key[c] ? k2 : key[c];
            }
        }
        for (int c = 0; c < 4; c++) red[c][tid] = key[c];
        elas_lds_barrier();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s)
                for (int c = 0; c < 4; c++) red[c][tid] = red[c][tid + s] < red[c][tid] ? red[c][tid + s] : red[c][tid];
            elas_lds_barrier();
        }
        if (tid < 6) {
            const int c = tid < 4 ? tid : tid - 2;  // the last two repeat the right corners
            const int i = (int)(unsigned)red[c][0];
            const int d = (int)P.sp[0][2 * i] - (int)P.sp[1][2 * i];
            const int u = tid < 4 ? cu[c] : P.W - 1 + d, v = cv[c];
            const float2 a = make_float2((float)u, (float)v), b = make_float2((float)(u - d), (float)v);
            reinterpret_cast<float2*>(P.sp[0])[n + tid] = a;
            reinterpret_cast<float2*>(P.sp[1])[n + tid] = b;
            reinterpret_cast<float2*>(P.hsp[0])[n + tid] = a;
            reinterpret_cast<float2*>(P.hsp[1])[n + tid] = b;
        }
        elas_lds_barrier();
        if (tid == 0) {
            P.nsp[0] = n + 6;
            P.hnsp[0] = n + 6;
        }
    }
}

void elas_launch_corners(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_corners, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_compact(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_compact, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_grid_fill(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_grid_fill, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_grid_scatter(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_grid_scatter, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_grid_dilate(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_grid_dilate, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_grid_list(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_grid_list, dim3(grid), dim3(256), 0, s, L);
}
