// k_dense.hip -- plane prior, triangle rasterisation, dense matching and the
// left/right check (gfx950).  Compiled with -ffp-contract=off; the float
// plane and span arithmetic also uses explicit _rn intrinsics so no FMA is
// formed.
#include <hip/hip_runtime.h>
#include "elas_common.h"
#include "elas_device.h"

// ---- plane solve: kerSolve, KerComputeDisparity.cu:322-375 ---------------
// Gauss-Jordan with full pivoting in double (eps 1e-20), the solution
// overwriting B, operation for operation as the reference.  The reference
// swaps the pivot row into position icol; here rows stay in their registers
// and pos[r] is the position row r would hold, so every access is by a fixed
// index and only the pivot's row and column are picked by selects.  Its
// pivot search scans positions j then columns k and takes the last element
// with |a| >= the running maximum: the largest (|a|, j, k), which is what the
// search below takes over the live (row, column) pairs.
__device__ bool elas_solve3(double A[3][3], double B[3]) {
    const double eps = 1e-20;
    int pos[3] = {0, 1, 2};
    bool used[3] = {false, false, false}, cdone[3] = {false, false, false};
    int pr = 0, pc = 0;  // pivot row (register index) and column
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double big = -1.0;
        int bj = -1, bk = -1;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double a = fabs(A[r][k]);
                const bool later = pos[r] > bj || (pos[r] == bj && k > bk);
                if (!used[r] && !cdone[k] && (a > big || (a == big && later))) {
                    big = a;
                    bj = pos[r];
                    bk = k;
                    pr = r;
                    pc = k;
                }
            }
        cdone[pc] = true;
        // the row at position pc takes the pivot row's old position
        const int irow = pos[pr];
#pragma unroll
        for (int r = 0; r < 3; r++)
            if (pos[r] == pc) pos[r] = irow;
        pos[pr] = pc;
        double piv = 0.0, prow[3], pb = 0.0;
#pragma unroll
        for (int r = 0; r < 3; r++)
            if (r == pr) {
#pragma unroll
                for (int l = 0; l < 3; l++) prow[l] = A[r][l];
                pb = B[r];
            }
#pragma unroll
        for (int l = 0; l < 3; l++)
            if (l == pc) piv = prow[l];
        if (fabs(piv) < eps) return false;
        const double pivinv = 1.0 / piv;
#pragma unroll
        for (int l = 0; l < 3; l++) prow[l] = __dmul_rn(l == pc ? 1.0 : prow[l], pivinv);
        pb = __dmul_rn(pb, pivinv);
#pragma unroll
        for (int r = 0; r < 3; r++) {
            double dum = 0.0;
#pragma unroll
            for (int l = 0; l < 3; l++)
                if (l == pc) dum = A[r][l];
            const bool isp = r == pr;
#pragma unroll
            for (int l = 0; l < 3; l++) {
                const double e = __dsub_rn(l == pc ? 0.0 : A[r][l], __dmul_rn(prow[l], dum));
                A[r][l] = isp ? prow[l] : e;
            }
            const double eb = __dsub_rn(B[r], __dmul_rn(pb, dum));
            B[r] = isp ? pb : eb;
        }
        used[pr] = true;
    }
    // the reference's B holds the solution by position: variable k at
    // position k
    double x[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int r = 0; r < 3; r++)
            if (pos[r] == k) x[k] = B[r];
#pragma unroll
    for (int k = 0; k < 3; k++) B[k] = x[k];
    return true;
}

// Plane coefficients and sorted-corner edge lines of triangle ti of image img.
__device__ __forceinline__ void elas_plane_rec(const ElasPair& P, int img, int ti, float* rec) {
    // corners: 16-bit indices (meshes of fewer than 65536 points) or 32-bit
    const int32_t* tr = P.tri[img] + 3 * (size_t)ti;
    const uint16_t* tr16 = (const uint16_t*)P.tri[img] + 3 * (size_t)ti;
    const float* sp = P.sp[0];
    const float* sq = P.sp[1];
    int tu[3], tv[3], td[3];
    double A[3][3], b[3];
    int ci[3];
    if (P.tri16)
        for (int i = 0; i < 3; i++) ci[i] = tr16[i];
    else
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

// ---- per-triangle record and culling masks: KerComputeDisparity.cu:10-124 --
// A 64-lane workgroup takes 64 consecutive triangles of one mesh (chunk c of
// image img), one lane each.  Corners come from the left support points sp0
// for both meshes; the right mesh shifts them by -d (Elas.cu:105,
// KerComputeDisparity.cu:63-66,95-99).  A corner's d is sp0.u - sp1.u (exact
// in float), which saves a dependent load of the lattice.
//
// The record carries the triangle's pixel box: columns [max(A_u, 0),
// min(C_u, W)) (the spans are half-open between corner columns) and the
// corner rows widened by a margin that covers the float error of the span
// ends (rows < 2^14, so both ends pack into one word).  A span end is the truncation of s*u + b, where
// s = fl((v_a - v_b) / (u_a - u_b)) and b = fl(v_a - fl(s*u_a)) for integer
// corners, evaluated at integer u between them as fl(fl(s*u) + b).  Against
// the exact line through the corners this is off by at most
// 2^-24 * (5*|s|*U + 2*V) (plus second-order terms), U and V bounding |u| and
// |v|; the margin is 3 + floor of 8 * (3*smax*U + 3*V) * 2^-24 (a float
// evaluation of a larger bound), so the span rows stay within
// [min v - margin, max v + margin].  The box then marks, in LDS, every
// super-tile it meets (a square of 1 << sshift raster tiles a side: a single
// tile unless the image has more than ELAS_ST_MAX tiles); every nonzero mask
// of the chunk is appended to its super-tile's list as (c, mask), bit k for
// triangle 64*c + k (the list order is immaterial: the raster takes a max).
// A list holds scap entries; the count goes on, and a tile whose list
// overflowed tests every triangle of the mesh instead.
__global__ __launch_bounds__(64) void k_plane(ElasLaunch L) {
    extern __shared__ unsigned long long sm[];  // [nst] of the largest pair of the launch
    __shared__ int32_t nzs[64];                 // a batch of super-tiles with nonzero masks
    const int lane = threadIdx.x;
    // the triangle counts are data: read the total from the table (a captured
    // launch keeps the grid of its shape bucket)
    const int total = L.prefix[L.npairs];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int nc0 = (P.ntri[0] + 63) >> 6;
        int c = t - L.prefix[p];
        const int img = c >= nc0;
        if (img) c -= nc0;
        const int nst = P.nst, nsx = P.nsx, sh = P.sshift;
        for (int s = lane; s < nst; s += 64) sm[s] = 0ull;
        __syncthreads();
        const int ti = c * 64 + lane;
        if (ti < P.ntri[img]) {
            float rec[ELAS_TRIREC];
            elas_plane_rec(P, img, ti, rec);
            const float Au = rec[3], Av = rec[4], Bv = rec[6], Cu = rec[7], Cv = rec[8];
            const float smax = fmaxf(fabsf(rec[9]), fmaxf(fabsf(rec[11]), fabsf(rec[13])));
            const float U = fmaxf(fabsf(Au), fabsf(Cu)) + 1.0f;
            const float V = fmaxf(fabsf(Av), fmaxf(fabsf(Bv), fabsf(Cv))) + 1.0f;
            const float x = 8.0f * (3.0f * smax * U + 3.0f * V) * (1.0f / 16777216.0f);
            const int margin = 3 + (int)fminf(x * 1.001f, 1.0e6f);
            const int vmin = (int)fminf(Av, fminf(Bv, Cv)), vmax = (int)fmaxf(Av, fmaxf(Bv, Cv));
            int4 bx = make_int4(max((int)Au, 0), min((int)Cu, P.W), max(vmin - margin, 0), min(vmax + margin + 1, P.H));
            if (bx.x >= bx.y || bx.z >= bx.w) bx = make_int4(0, 0, 0, 0);
            // the raster's record: (A_u, B_u, C_u, box rows v0 | v1 << 16),
            // the AB and AC edges, the BC edge, the plane
            float4* dst = (float4*)(P.trirec[img] + (size_t)ti * ELAS_TRIREC);
            dst[0] = make_float4(rec[3], rec[5], rec[7], __int_as_float(bx.z | (bx.w << 16)));
            dst[1] = make_float4(rec[9], rec[10], rec[11], rec[12]);
            dst[2] = make_float4(rec[13], rec[14], 0.f, 0.f);
            dst[3] = make_float4(rec[0], rec[1], rec[2], 0.f);
            if (bx.x < bx.y) {
                const int sx0 = bx.x >> (ELAS_RT_WS + sh), sx1 = (bx.y - 1) >> (ELAS_RT_WS + sh);
                const int sy0 = bx.z >> (ELAS_RT_HS + sh), sy1 = (bx.w - 1) >> (ELAS_RT_HS + sh);
                for (int sy = sy0; sy <= sy1; sy++)
                    for (int sx = sx0; sx <= sx1; sx++) atomicOr(&sm[sy * nsx + sx], 1ull << lane);
            }
        }
        __syncthreads();
        // appends: the nonzero masks are compacted into nzs, 64 at a time, and
        // each batch takes its list slots in one round trip (the workgroup is
        // one wave: its barriers only order the LDS accesses)
        auto flush = [&](int n) {
            __syncthreads();
            const int ms = lane < n ? nzs[lane] : 0;
            const unsigned long long m = lane < n ? sm[ms] : 0ull;
            const int slot = lane < n ? atomicAdd(P.scnt[img] + ms, 1) : 0;
            if (lane < n && slot < P.scap)
                ((int4*)P.slist[img])[(size_t)ms * P.scap + slot] =
                    make_int4(c, (int)(uint32_t)m, (int)(uint32_t)(m >> 32), 0);
            __syncthreads();
        };
        int filled = 0;
        for (int s0 = 0; s0 < nst; s0 += 64) {
            const int s = s0 + lane;
            const bool nz = s < nst && sm[s] != 0ull;
            const unsigned long long bal = __ballot(nz);
            const int nb = __popcll(bal);
            if (filled + nb > 64) {
                flush(filled);
                filled = 0;
            }
            if (nz) nzs[filled + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0))] = s;
            filled += nb;
        }
        if (filled) flush(filled);
        __syncthreads();
    }
}

// ---- rasterisation: KerComputeDisparity.cu:126-150 ------------------------
// A workgroup owns an ELAS_RT_W x ELAS_RT_H pixel tile of one image.  It
// reads its super-tile's list of 64-triangle chunks into LDS and gathers the
// listed triangles four chunks per round, a wave per chunk and a lane per
// triangle: a lane whose mask bit is set loads the triangle's record, tests
// its box against the tile and, on a hit, appends the walk's part of it to
// LDS (the edges, the triangle index and its columns in the tile).  The next
// round's records are loaded before this round's walk.  Wave w
// then takes the gathered triangles w, w + 4, ..., four at a time, one per
// 16-lane group; a group's lanes walk the tile columns u of the triangle's
// half-open column range [A_u, C_u), 16 apart: the column's row span between
// the long edge AC and AB (u < B_u) or BC, as the reference computes it,
// clipped to the tile.
// Each pixel of the span takes the LDS atomic max of the triangle index:
// where spans overlap (the reference races there) the highest index wins,
// the result of painting the triangles in index order.  Last, every pixel of
// the tile is written once: the plane prior (int)(a*u + b*v + c) of its
// triangle, or -1 where no span covers it.  No global atomics.
#define ELAS_RT_CH 64  // list entries (chunks) read into LDS per pass
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_rtile(ElasLaunch L) {
    __shared__ int32_t ids[ELAS_RT_H * ELAS_RT_W];
    // gathered triangle s: srec[2s] = (ABa, ABb, ACa, ACb), srec[2s+1] =
    // (BCa, BCb, columns, index); columns = ulo | uhi << 8 | bu << 16, the
    // triangle's columns [ulo, uhi) and B_u clamped to [0, 64], tile-relative
    __shared__ float4 srec[2 * 256];
    __shared__ unsigned long long nzm[ELAS_RT_CH];
    __shared__ int32_t nzc[ELAS_RT_CH];
    __shared__ int32_t nh[2];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int W = P.W, H = P.H;
        const int ntx = (W + ELAS_RT_W - 1) / ELAS_RT_W, nty = (H + ELAS_RT_H - 1) / ELAS_RT_H;
        int k = t - L.prefix[p];
        const int img = k >= ntx * nty;
        if (img) k -= ntx * nty;
        const int ty = k / ntx, tx = k - ty * ntx;
        const int u0 = tx * ELAS_RT_W, v0 = ty * ELAS_RT_H;
        const int sh = P.sshift;
        const int st = (tx >> sh) + (ty >> sh) * P.nsx;
        for (int i = tid; i < ELAS_RT_H * ELAS_RT_W; i += 256) ids[i] = -1;
        const int4* slist = (const int4*)P.slist[img] + (size_t)st * P.scap;
        // the count and the list's first entries in one round trip
        const int nlist = P.scnt[img][st];
        const int4 e0 = tid < ELAS_RT_CH && tid < P.scap ? slist[tid] : make_int4(0, 0, 0, 0);
        const int ntri = P.ntri[img], nch = (ntri + 63) >> 6;
        const bool over = nlist > P.scap;  // incomplete list: every chunk, every triangle
        const int nent = over ? nch : nlist;
        const float4* rec4 = (const float4*)P.trirec[img];
        for (int cb = 0; cb < nent; cb += ELAS_RT_CH) {
            const int j = cb + tid;
            if (tid < ELAS_RT_CH && j < nent) {
                if (over) {
                    nzc[tid] = j;
                    nzm[tid] = ntri - 64 * j >= 64 ? ~0ull : (1ull << (ntri - 64 * j)) - 1ull;
                } else {
                    const int4 e = j < ELAS_RT_CH ? e0 : slist[j];
                    nzc[tid] = e.x;
                    nzm[tid] = (unsigned long long)(uint32_t)e.y | ((unsigned long long)(uint32_t)e.z << 32);
                }
            }
            if (tid == 0) nh[0] = 0;
            __syncthreads();
            const int nz = min(ELAS_RT_CH, nent - cb);
            // a round's records: wave w takes chunk jq = r + w, a lane per
            // triangle; the next round's go out before this round's walk
            int ti;
            bool in;
            float4 a0, a1, a2;
            auto fetch = [&](int jq) {
                ti = (jq < nz ? nzc[jq] : 0) * 64 + lane;
                in = jq < nz && ((nzm[jq] >> lane) & 1ull);
                a0 = make_float4(0.f, 0.f, 0.f, 0.f);
                a1 = a2 = a0;
                if (in) {
                    a0 = rec4[(size_t)ti * 4];
                    a1 = rec4[(size_t)ti * 4 + 1];
                    a2 = rec4[(size_t)ti * 4 + 2];
                }
            };
            fetch(wave);
            for (int r = 0; r < nz; r += 4) {
                // round r appends to nh[h]; the other count is cleared for the
                // next round once every wave has read it
                const int h = (r >> 2) & 1;
                const int rows = __float_as_int(a0.w);
                const int ulo = max(max((int)a0.x, 0), u0), uhi = min(min((int)a0.z, W), u0 + ELAS_RT_W);
                const bool hit = in && ulo < uhi && (rows & 0xffff) < v0 + ELAS_RT_H && (rows >> 16) > v0;
                const unsigned long long bal = __ballot(hit);
                if (bal) {
                    int base = 0;
                    if (lane == 0) base = atomicAdd(&nh[h], __popcll(bal));
                    base = __builtin_amdgcn_readfirstlane(base);
                    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0));
                    if (hit) {
                        const int bu = min(max((int)a0.y - u0, 0), ELAS_RT_W);
                        const int cols = (ulo - u0) | ((uhi - u0) << 8) | (bu << 16);
                        srec[2 * (base + below)] = a1;
                        srec[2 * (base + below) + 1] = make_float4(a2.x, a2.y, __int_as_float(cols), __int_as_float(ti));
                    }
                }
                __syncthreads();
                const int n = nh[h];
                if (tid == 0) nh[h ^ 1] = 0;
                fetch(r + 4 + wave);
                const int cnt = n > wave ? (n - wave + 3) >> 2 : 0;
                // the wave's triangles w + 4 i, four at a time: lane group q
                // (16 lanes) takes triangle g + q and walks its columns in the
                // tile, lane c the columns first + c + 16 k
                const int q = lane >> 4, c = lane & 15;
                for (int g = 0; g < cnt; g += 4) {
                    if (g + q < cnt) {
                        const int sl = wave + 4 * (g + q);
                        const float4 r0 = srec[2 * sl], r1 = srec[2 * sl + 1];
                        const float ABa = r0.x, ABb = r0.y, ACa = r0.z, ACb = r0.w, BCa = r1.x, BCb = r1.y;
                        const int cols = __float_as_int(r1.z), tr = __float_as_int(r1.w);
                        const int uhi = (cols >> 8) & 255, bu = cols >> 16;
                        for (int uo = (cols & 255) + c; uo < uhi; uo += 16) {
                            const float fu = (float)(u0 + uo);
                            const bool first = uo < bu;
                            const float ea = first ? ABa : BCa, eb = first ? ABb : BCb;
                            const int v1 = (int)elas_f2u_sat(__fadd_rn(__fmul_rn(ACa, fu), ACb));
                            const int v2 = (int)elas_f2u_sat(__fadd_rn(__fmul_rn(ea, fu), eb));
                            const int vlo = max(max(min(v1, v2), 0), v0);
                            const int vhi = min(min(max(v1, v2), H), v0 + ELAS_RT_H);
                            for (int v = vlo; v < vhi; v++) atomicMax(&ids[(v - v0) * ELAS_RT_W + uo], tr);
                        }
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
        // every pixel once: all of a thread's plane loads (one 16-byte load of
        // the record's a, b, c) go out before its stores
        int32_t* prior = P.prior[img];
        constexpr int PER = ELAS_RT_H * ELAS_RT_W / 256;
        int idv[PER];
        float4 pl[PER];
        #pragma unroll
        for (int k = 0; k < PER; k++) {
            idv[k] = ids[tid + 256 * k];
            pl[k] = rec4[(size_t)(idv[k] >= 0 ? idv[k] : 0) * 4 + 3];
        }
        #pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + 256 * k;
            const int uu = u0 + (i % ELAS_RT_W), v = v0 + i / ELAS_RT_W;
            if (uu < W && v < H) {
                const int d = elas_f2i_sat(__fadd_rn(__fadd_rn(__fmul_rn(pl[k].x, (float)uu), __fmul_rn(pl[k].y, (float)v)), pl[k].z));
                prior[(size_t)v * W + uu] = idv[k] >= 0 ? d : -1;
            }
        }
        __syncthreads();
    }
}

// ---- phase B gate ------------------------------------------------------------
// Queued ahead of phase B while the host triangulates: lane 0 polls the
// release flag in mapped pinned memory (system-scope acquire loads, about
// 4 us apart) for at most `limit` ticks of the 100 MHz wall clock (3e9, 30 s,
// unless a test shortens it), then the workgroup copies the phase B table
// into the device table: once released, `ready`, which the host wrote into
// pinned memory before setting the flag; after a timeout, `fallback`, a table
// written before the launch and left alone while the gate waits (the host may
// be writing `ready` at that moment), and status[0] tells the host.
__global__ __launch_bounds__(256) void k_gate(const uint32_t* flag, uint32_t* status, const uint32_t* ready,
                                              const uint32_t* fallback, uint32_t* dst, int words, uint64_t limit) {
    __shared__ uint32_t released;
    if (threadIdx.x == 0) {
        const uint64_t t0 = wall_clock64();
        uint32_t ok = 0u;
        for (;;) {
            if (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
                ok = 1u;
                break;
            }
            if (wall_clock64() - t0 > limit) break;
            // about 4 us between polls: each poll is a coherent read of host
            // memory, and polling much faster slows the host's Triangle
            __builtin_amdgcn_s_sleep(127);
        }
        if (!ok) __hip_atomic_store(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        released = ok;
    }
    __syncthreads();
    const uint32_t* src = released ? ready : fallback;
    for (int i = threadIdx.x; i < words; i += 256)
        dst[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
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
// Work unit = (pair, band, chunk).  A band is R = min(64/gs, gs) image rows
// inside one grid-cell row, matched for both images at once; a chunk is a run
// of up to P.dspc steps along the band, a step covering tw = WAVES*P.dcpw*gs
// columns.  In a step, wave w owns cells s*WAVES*dcpw + w + WAVES*j and a lane
// is a (row, column) of a cell, so the cell's candidate list is wave-uniform:
// the grid kernel wrote it as 16-bit entries, a wave loads a step's lists
// into lanes of one register (word k of a list in lane k of its share) and
// reads two candidates at a time into a scalar register.
//
// LDS holds two rings of R descriptor rows:
//   ring 0: desc0 over [u_lo, u_hi + dmax): the left pixels' own blocks and
//           the right pixels' matches at u + d;
//   ring 1: desc1 over [u_lo - dmax, u_hi): the right pixels' own blocks and
//           the left pixels' matches at u - d.
// A ring row has RW slots, the smallest power of two >= tw + dmax, and image
// column c sits in slot c mod RW of its row, so the tw columns a step brings
// in take the slots of columns that have left, and each descriptor row of a
// chunk is read from HBM once.  Row l of the rings is rotated by 4 l slots so
// that a wave's rows start on different banks: column c of line l is at byte
// 16 RW l | ((16 c + 64 l) mod 16 RW), one add and one and-or per address.
// The next step's columns, priors and candidate lists are loaded into
// registers while the current step matches and are written after it.
//
// Reference order and tie rule: grid candidates outside the plane window
// [dp-r, dp+r] ascending, then the window ascending with the prior P added,
// strict '<'.  This equals the minimum of the key
//   ((cost + bias) << 16) | (phase << 10) | x,   phase 0 = grid, x = d;
//                                                phase 1 = window, x = d - dp + r,
// built by v_sad_hi_u8, which adds each byte block's SAD << 16 onto the
// candidate's constant ((bias [+ P]) << 16 | phase << 10 | x).  cost + bias
// stays below 2^15 (bias <= ELAS_MAX_KEY_BIAS), and a candidate the reference
// does not evaluate starts from KEY_BAD = 2^31, above every valid key, with
// room for the SAD below 2^32.  When every P(k) < 0, a grid candidate inside
// the window is dominated by its own window entry (same d, lower cost), so
// phase 0 need not exclude the window; otherwise it does.

// smallest power of two >= x (x >= 1)
__device__ __forceinline__ uint32_t elas_pow2_at_least(uint32_t x) { return x <= 1u ? 1u : 1u << (32 - __clz(x - 1u)); }

// Workgroups are dispatched to the 8 XCDs round robin (workgroup b runs on XCD
// b mod 8).  With a grid that is a multiple of 8, work item t of a grid-stride
// loop runs on XCD t mod 8; this order gives each XCD a contiguous run of the
// work list, so neighbouring units (which share descriptor columns) meet in
// the same L2.  A bijection of [0, total).
__device__ __forceinline__ int elas_xcd_order(int t, int total, int grid) {
    if (grid & 7) return t;
    const int x = t & 7, i = t >> 3, q = total >> 3, r = total & 7;
    return x * q + min(x, r) + i;
}

template <int N>
struct elas_ic {
    static constexpr int value = N;
};

// CPW = P.dcpw: cells per wave per step (a launch holds the pairs of one CPW).
// Global loads and stores are buffer accesses issued without a branch (lanes
// with nothing to move take an out-of-range offset), so the waits below are
// for exactly the loads they need.  Ring slots of columns outside [3, W-4]
// hold whatever the loads returned: no valid key reads them (a pixel is
// matched only for u in [4, W-4] and u -/+ d in [3, W-4]).
template <int CPW, bool PNEG, bool TEX>
__global__ __launch_bounds__(ELAS_DENSE_THREADS, CPW == 1 ? 6 : 4) void k_dense(ElasLaunch L, ElasKParams kp) {
    extern __shared__ uint4 s_ring[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar registers
    constexpr uint32_t KEY_BAD = 0x80000000u;
    // lanes per candidate list in the list register: HC words of list h in
    // lanes [HC h, HC h + HC), h = img * CPW + ci
    constexpr int HC = 64 / (2 * CPW);
    // ring reads by LDS byte address: the dynamic LDS (the rings) starts
    // right after the kernel's static LDS, of which it has none, so at 0
    typedef const __attribute__((address_space(3))) elas_u32x4 lds_u4;
    auto ld = [&](uint32_t a) {
        const elas_u32x4 x = *(lds_u4*)(uintptr_t)a;
        return make_uint4(x.x, x.y, x.z, x.w);
    };
    const int pr = kp.plane_radius, nwin = 2 * pr + 1;
    const uint32_t kgrid = (uint32_t)kp.key_bias << 16;
    for (int t0 = blockIdx.x; t0 < L.total; t0 += gridDim.x) {
        const int t = elas_xcd_order(t0, L.total, gridDim.x);
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int W = P.W, H = P.H, gs = P.gsize, gw = P.gw, dmax = P.dmax;
        const int R = max(1, min(64 / gs, gs));
        const int sub = (gs + R - 1) / R;
        const int tcells = ELAS_DENSE_WAVES * CPW, tw = tcells * gs;
        const int segs = (gw + tcells - 1) / tcells;
        const int chunks = (segs + P.dspc - 1) / P.dspc;
        const int r = t - L.prefix[p];
        const int band = r / chunks, chunk = r - band * chunks;
        const int gy = band / sub, kb = band - gy * sub;
        const int v0 = gy * gs + kb * R;
        const int rows = max(0, min(min(R, gs - kb * R), H - v0));
        if (rows == 0) continue;  // workgroup-uniform
        const int s0 = chunk * P.dspc, s1 = min(segs, s0 + P.dspc);
        const int RW = (int)elas_pow2_at_least((uint32_t)(tw + dmax));
        const uint32_t M16 = 16u * (uint32_t)RW - 1u;
        // uint4 index of column c in line l
        auto phys = [&](int line, int c) { return line * RW + ((c + 4 * line) & (RW - 1)); };
        const uint32_t npx = (uint32_t)W * (uint32_t)H;  // <= 2^27
        const __amdgpu_buffer_rsrc_t rd0 = elas_rsrc(P.desc[0], 16u * npx);
        const __amdgpu_buffer_rsrc_t rd1 = elas_rsrc(P.desc[1], 16u * npx);
        const int row = lane / gs, col = lane - row * gs;
        const int lr = row < rows ? row : 0;
        const int v = v0 + row;
        const bool lane_on = lane < R * gs && row < rows;
        const bool row_ok = lane_on && v >= 4 && v <= H - 4;
        const uint32_t vrow4 = 4u * (uint32_t)v * (uint32_t)W;  // byte offset of this lane's row in a float/int map
        const __amdgpu_buffer_rsrc_t rp[2] = {elas_rsrc(P.prior[0], 4u * npx), elas_rsrc(P.prior[1], 4u * npx)};
        const uint32_t cwords = (uint32_t)P.cstride >> 1;  // 32-bit words of a cell's list
        const uint32_t lbytes = 4u * (uint32_t)(P.gh * gw) * cwords;  // one image's lists
        const __amdgpu_buffer_rsrc_t rc = elas_rsrc(P.gcand[0], 2u * lbytes);
        const __amdgpu_buffer_rsrc_t ro[2] = {elas_rsrc(P.draw[0], 4u * npx), elas_rsrc(P.draw[1], 4u * npx)};
        // priors of this lane's pixel in each (image, cell) of step s, and the
        // step's candidate lists (lane HC h + k: word k of list h)
        auto load_px = [&](int s, int (&prv)[2][CPW], uint32_t& lst) {
            {
                const int h = lane / HC, w = lane - h * HC, img = h / CPW, ci = h - img * CPW;
                const int gx = s * tcells + wave + ci * ELAS_DENSE_WAVES;
                const bool ok = gx < gw && s < s1 && w < (int)cwords;
                lst = elas_ld4_ranged(rc, ok ? (uint32_t)img * lbytes + ((uint32_t)(gy * gw + gx) * cwords + (uint32_t)w) * 4u
                                             : ELAS_OOR);
            }
            #pragma unroll
            for (int c = 0; c < CPW; c++) {
                const int gx = s * tcells + wave + c * ELAS_DENSE_WAVES;
                const uint32_t off = gx < gw && s < s1 ? vrow4 + 4u * (uint32_t)(gx * gs + col) : ELAS_OOR;
                #pragma unroll
                for (int img = 0; img < 2; img++) prv[img][c] = (int)elas_ld4_ranged(rp[img], off);
            }
        };
        // this step's priors and lists are in flight with the ring loads below
        int pr_cur[2][CPW];
        uint32_t ls_cur;
        load_px(s0, pr_cur, ls_cur);
        // prologue: step s0's windows, both rings at the same (row, j), every
        // load of a batch issued before its stores
        {
            auto dload = [&](int k, int rr, int c, bool ok) {
                const uint32_t off =
                    ok && c >= 0 && c < W ? ((uint32_t)(v0 + rr) * (uint32_t)W + (uint32_t)c) * 16u : ELAS_OOR;
                return elas_ld16_ranged(k ? rd1 : rd0, off);
            };
            const int span = tw + dmax, n = rows * span, u_lo = s0 * tw;
            int rr = tid / span, j = tid - (tid / span) * span;
            for (int base = 0; base < n; base += 2 * ELAS_DENSE_THREADS) {
                uint4 q[2][2];
                int li[2];
                #pragma unroll
                for (int e = 0; e < 2; e++) {
                    const bool ok = base + tid + e * ELAS_DENSE_THREADS < n;
                    q[0][e] = dload(0, rr, u_lo + j, ok);
                    q[1][e] = dload(1, rr, u_lo - dmax + j, ok);
                    li[e] = ok ? (rr << 16) | j : -1;
                    j += ELAS_DENSE_THREADS;
                    while (j >= span) {
                        j -= span;
                        rr++;
                    }
                }
                #pragma unroll
                for (int e = 0; e < 2; e++)
                    if (li[e] >= 0) {
                        const int lrr = li[e] >> 16, jj = li[e] & 0xffff;
                        s_ring[phys(lrr, u_lo + jj)] = q[0][e];
                        s_ring[phys(rows + lrr, u_lo - dmax + jj)] = q[1][e];
                    }
            }
        }
        // this thread's share of a step's new columns, the same (row, j) in
        // both rings (rows x tw <= 64 * 8 * CPW per ring): the ring's row and
        // column offset of step s + 1's j-th new column, ring 0 at
        // (s + 1) tw + dmax + j, ring 1 at (s + 1) tw + j; byte offsets of
        // those columns in the descriptor images at s + 1 = 0
        int pf_line[CPW], pf_j[CPW];
        uint32_t pf_off[CPW];
        #pragma unroll
        for (int e = 0; e < CPW; e++) {
            const int i = tid + e * ELAS_DENSE_THREADS;
            const bool has = i < rows * tw;
            pf_line[e] = has ? i / tw : -1;
            pf_j[e] = has ? i - (i / tw) * tw : 0;
            pf_off[e] = has ? ((uint32_t)(v0 + pf_line[e]) * (uint32_t)W + (uint32_t)pf_j[e]) * 16u : ELAS_OOR;
        }
        elas_lds_barrier();
        for (int s = s0; s < s1; s++) {
            // next step's new columns, priors and lists; past the chunk's last
            // step these loads read nothing
            const bool more = s + 1 < s1;
            const int u_nx = (s + 1) * tw;
            uint4 pf[2][CPW];
            #pragma unroll
            for (int e = 0; e < CPW; e++) {
                const bool ok = more && pf_line[e] >= 0;
                pf[0][e] = elas_ld16_ranged(rd0, ok ? pf_off[e] + 16u * (uint32_t)(u_nx + dmax) : ELAS_OOR);
                pf[1][e] = elas_ld16_ranged(rd1, ok ? pf_off[e] + 16u * (uint32_t)u_nx : ELAS_OOR);
            }
            int pr_nx[2][CPW];
            uint32_t ls_nx;
            load_px(s + 1, pr_nx, ls_nx);
            #pragma unroll
            for (int ci = 0; ci < CPW; ci++) {
                const int gx = s * tcells + wave + ci * ELAS_DENSE_WAVES;
                if (gx >= gw) continue;  // wave-uniform
                const int u = gx * gs + col;
                const bool active = lane_on && u < W;
                const bool col_ok = row_ok && u >= 4 && u <= W - 4;
                const uint32_t u16 = 16u * (uint32_t)u;
                #pragma unroll
                for (int img = 0; img < 2; img++) {
                    const int h = img * CPW + ci;
                    const int dp = pr_cur[img][ci];
                    const int own_line = img * rows + lr, oth_line = (1 - img) * rows + lr;
                    const uint32_t ob = 16u * (uint32_t)(RW * oth_line);
                    // own block; the match at disparity d is column u -/+ d of the other ring
                    const uint4 b1 = ld((16u * (uint32_t)(RW * own_line)) | ((u16 + 64u * (uint32_t)own_line) & M16));
                    const uint32_t xo = u16 + 64u * (uint32_t)oth_line;
                    bool valid = col_ok && dp > 0;
                    if (TEX) {  // match_texture > 0
                        const uint32_t tex =
                            elas_sad16(b1, make_uint4(0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u), 0u);
                        valid = valid && (int)tex >= kp.match_texture;
                    }
                    // u_warp = u -/+ d must lie in [3, W-4] (KerComputeDisparity.cu:259,267): d <= dlim
                    const int dlim = img ? W - 4 - u : u - 3;
                    const int wlim = min(dmax, dlim);
                    uint32_t kmin = 0xffffffffu;
                    // plane window d = dp - r + q, q = 0..2r
                    const uint32_t xw = img ? xo + 16u * (uint32_t)(dp - pr) : xo - 16u * (uint32_t)(dp - pr);
                    auto waddr = [&](int q) {
                        return ob | ((img ? xw + 16u * (uint32_t)q : xw - 16u * (uint32_t)q) & M16);
                    };
                    // every valid lane's window inside [0, wlim] (the usual case): no per-lane test
                    const bool win_in = __ballot(valid && !(dp >= pr && dp + pr <= wlim)) == 0;
                    // window keys: ((bias + P(|q - r|)) << 16) | 1024 | q, plus the SAD;
                    // an entry off [0, wlim] starts from KEY_BAD (its wrapped,
                    // in-bounds read need not be clamped)
                    const uint32_t bias = (uint32_t)kp.key_bias;
                    auto window = [&]() {
                        if (nwin == 5) {
                            // the effective parameters: the priors P(2), P(1), P(0), P(1),
                            // P(2) indexed by constants (a runtime index into the kernel
                            // arguments is a scalar memory load per entry)
                            const uint32_t kw[5] = {((bias + kp.P[2]) << 16) | 1024u, ((bias + kp.P[1]) << 16) | 1025u,
                                                    ((bias + kp.P[0]) << 16) | 1026u, ((bias + kp.P[1]) << 16) | 1027u,
                                                    ((bias + kp.P[2]) << 16) | 1028u};
                            uint4 wb[5];
                            #pragma unroll
                            for (int q = 0; q < 5; q++) wb[q] = ld(waddr(q));
                            if (win_in) {
                                #pragma unroll
                                for (int q = 0; q < 5; q++) kmin = min(kmin, elas_sad16_hi(b1, wb[q], kw[q]));
                            } else {
                                #pragma unroll
                                for (int q = 0; q < 5; q++) {
                                    const bool ok = (uint32_t)(dp - 2 + q) <= (uint32_t)wlim;
                                    kmin = min(kmin, elas_sad16_hi(b1, wb[q], ok ? kw[q] : KEY_BAD));
                                }
                            }
                        } else {
                            for (int q = 0; q < nwin; q++) {
                                const bool ok = (uint32_t)(dp - pr + q) <= (uint32_t)wlim;
                                const uint32_t c = ((bias + (uint32_t)kp.P[abs(q - pr)]) << 16) | 1024u | (uint32_t)q;
                                kmin = min(kmin, elas_sad16_hi(b1, ld(waddr(q)), ok ? c : KEY_BAD));
                            }
                        }
                    };
                    // grid candidates: entries k of list h, two per list word, in
                    // batches whose reads are all issued before their SADs; the
                    // first batch's reads go out with the window's
                    const uint32_t w0 = __builtin_amdgcn_readlane(ls_cur, HC * h);
                    const int n = (int)(w0 & 0xffffu), last = (int)(w0 >> 16);
                    const int dlim_min = img ? W - 4 - (gx * gs + gs - 1) : gx * gs - 3;
                    const bool fast = PNEG && last <= dlim_min;
                    const int dpmin = max(dp - pr, 0), dpmax = min(dp + pr, dmax);
                    auto word = [&](int kw) { return (uint32_t)__builtin_amdgcn_readlane(ls_cur, HC * h + kw); };
                    auto issue = [&](auto nb, const int* dq, uint4* b2) {
                        constexpr int NB = decltype(nb)::value;
                        #pragma unroll
                        for (int q = 0; q < NB; q++)
                            b2[q] = ld(ob | ((img ? xo + 16u * (uint32_t)dq[q] : xo - 16u * (uint32_t)dq[q]) & M16));
                    };
                    auto consume = [&](auto nb, const int* dq, const uint4* b2) {
                        constexpr int NB = decltype(nb)::value;
                        if (fast) {
                            #pragma unroll
                            for (int q = 0; q < NB; q++) kmin = min(kmin, elas_sad16_hi(b1, b2[q], kgrid | (uint32_t)dq[q]));
                        } else {
                            #pragma unroll
                            for (int q = 0; q < NB; q++) {
                                bool ok = dq[q] <= dlim;
                                if (!PNEG) ok = ok && (dq[q] < dpmin || dq[q] > dpmax);
                                kmin = min(kmin, elas_sad16_hi(b1, b2[q], ok ? kgrid | (uint32_t)dq[q] : KEY_BAD));
                            }
                        }
                    };
                    auto batch = [&](auto nb, const int* dq) {
                        uint4 b2[decltype(nb)::value];
                        issue(nb, dq, b2);
                        consume(nb, dq, b2);
                    };
                    int k;
                    if (n >= 4) {
                        const uint32_t wa = word(1), wb2 = word(2);
                        const int dq[4] = {(int)(wa & 0xffffu), (int)(wa >> 16), (int)(wb2 & 0xffffu), (int)(wb2 >> 16)};
                        uint4 b2[4];
                        issue(elas_ic<4>(), dq, b2);
                        window();
                        consume(elas_ic<4>(), dq, b2);
                        k = 4;
                    } else if (n >= 2) {
                        const uint32_t wa = word(1);
                        const int dq[2] = {(int)(wa & 0xffffu), (int)(wa >> 16)};
                        uint4 b2[2];
                        issue(elas_ic<2>(), dq, b2);
                        window();
                        consume(elas_ic<2>(), dq, b2);
                        k = 2;
                    } else {
                        window();
                        k = 0;
                    }
                    // entries in the list register, then (lists longer than 2 HC - 2
                    // entries) the rest from memory
                    const int nr = min(n, 2 * HC - 2);
                    for (; k + 4 <= nr; k += 4) {
                        const uint32_t wa = word(1 + (k >> 1)), wb2 = word(2 + (k >> 1));
                        const int dq[4] = {(int)(wa & 0xffffu), (int)(wa >> 16), (int)(wb2 & 0xffffu), (int)(wb2 >> 16)};
                        batch(elas_ic<4>(), dq);
                    }
                    if (k + 2 <= nr) {
                        const uint32_t wa = word(1 + (k >> 1));
                        const int dq[2] = {(int)(wa & 0xffffu), (int)(wa >> 16)};
                        batch(elas_ic<2>(), dq);
                        k += 2;
                    }
                    if (k < nr) {
                        const int dq[1] = {(int)(word(1 + (k >> 1)) & 0xffffu)};
                        batch(elas_ic<1>(), dq);
                        k++;
                    }
                    if (k < n) {
                        const ELAS_CONST_AS uint32_t* gl =
                            (const ELAS_CONST_AS uint32_t*)P.gcand[img] + (size_t)(gy * gw + gx) * cwords;
                        for (; k < n; k++) {
                            const uint32_t wk = gl[1 + (k >> 1)];
                            const int dq[1] = {(int)((k & 1) ? wk >> 16 : wk & 0xffffu)};
                            batch(elas_ic<1>(), dq);
                        }
                    }
                    float res = -1.0f;
                    if (valid && kmin < KEY_BAD) {
                        const int x = (int)(kmin & 1023u);
                        res = (float)((kmin & 1024u) ? dp - pr + x : x);
                    }
                    elas_st4_ranged(ro[img], active ? vrow4 + 4u * (uint32_t)u : ELAS_OOR, __float_as_uint(res));
                }
            }
            if (more) {
                elas_lds_barrier();  // every wave is done with step s's slots
                #pragma unroll
                for (int e = 0; e < CPW; e++)
                    if (pf_line[e] >= 0) {
                        s_ring[phys(pf_line[e], u_nx + dmax + pf_j[e])] = pf[0][e];
                        s_ring[phys(rows + pf_line[e], u_nx + pf_j[e])] = pf[1][e];
                    }
                elas_lds_barrier();
            }
            #pragma unroll
            for (int img = 0; img < 2; img++)
                #pragma unroll
                for (int c = 0; c < CPW; c++) pr_cur[img][c] = pr_nx[img][c];
            ls_cur = ls_nx;
        }
        elas_lds_barrier();  // the next unit's prologue rewrites the rings
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

void elas_launch_plane(const ElasLaunch& L, int grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_plane, dim3(grid), dim3(64), lds, s, L);
}
void elas_launch_gate(const uint32_t* flag, uint32_t* status, const uint32_t* ready, const uint32_t* fallback,
                      uint32_t* dst, int words, uint64_t limit, hipStream_t s) {
    hipLaunchKernelGGL(k_gate, dim3(1), dim3(256), 0, s, flag, status, ready, fallback, dst, words, limit);
}
void elas_launch_rtile(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_rtile, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_mean(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_mean, dim3(grid), dim3(256), 0, s, L);
}
void elas_launch_dense(const ElasLaunch& L, const ElasKParams& kp, int cpw, int grid, size_t lds, hipStream_t s) {
    const dim3 b(ELAS_DENSE_THREADS);
    const bool pn = kp.all_p_negative != 0, tex = kp.match_texture > 0;
#define ELAS_DENSE_GO(C, PN, T) hipLaunchKernelGGL((k_dense<C, PN, T>), dim3(grid), b, lds, s, L, kp)
    const int which = (cpw == 2 ? 4 : 0) | (pn ? 2 : 0) | (tex ? 1 : 0);
    switch (which) {
        case 0: ELAS_DENSE_GO(1, false, false); break;
        case 1: ELAS_DENSE_GO(1, false, true); break;
        case 2: ELAS_DENSE_GO(1, true, false); break;
        case 3: ELAS_DENSE_GO(1, true, true); break;
        case 4: ELAS_DENSE_GO(2, false, false); break;
        case 5: ELAS_DENSE_GO(2, false, true); break;
        case 6: ELAS_DENSE_GO(2, true, false); break;
        default: ELAS_DENSE_GO(2, true, true); break;
    }
#undef ELAS_DENSE_GO
}
int elas_dense_allow_lds(size_t bytes) {
    const void* fn[8] = {reinterpret_cast<const void*>(&k_dense<1, true, false>),
                         reinterpret_cast<const void*>(&k_dense<1, false, false>),
                         reinterpret_cast<const void*>(&k_dense<2, true, false>),
                         reinterpret_cast<const void*>(&k_dense<2, false, false>),
                         reinterpret_cast<const void*>(&k_dense<1, true, true>),
                         reinterpret_cast<const void*>(&k_dense<1, false, true>),
                         reinterpret_cast<const void*>(&k_dense<2, true, true>),
                         reinterpret_cast<const void*>(&k_dense<2, false, true>)};
    for (const void* f : fn)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) {
            (void)hipGetLastError();
            return 1;
        }
    return 0;
}
void elas_launch_lr(const ElasLaunch& L, const ElasKParams& kp, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_lr, dim3(grid), dim3(256), 0, s, L, kp);
}
