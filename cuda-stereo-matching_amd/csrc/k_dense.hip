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

// ---- per-triangle record and culling masks: KerComputeDisparity.cu:10-124 --
// A 64-lane workgroup takes 64 consecutive triangles of one mesh (chunk c of
// image img), one lane each.  Corners come from the left support points sp0
// for both meshes; the right mesh shifts them by -d (Elas.cu:105,
// KerComputeDisparity.cu:63-66,95-99).  A corner's d is sp0.u - sp1.u (exact
// in float), which saves a dependent load of the lattice.
//
// Besides the record, a lane stores its triangle's pixel box: columns
// [max(A_u, 0), min(C_u, W)) (the spans are half-open between corner
// columns) and the corner rows widened by a margin that covers the float
// error of the span ends.  A span end is the truncation of s*u + b, where
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
            float4* dst = (float4*)(P.trirec[img] + (size_t)ti * ELAS_TRIREC);
            for (int k = 0; k < 4; k++) dst[k] = make_float4(rec[4 * k], rec[4 * k + 1], rec[4 * k + 2], rec[4 * k + 3]);
            const float Au = rec[3], Av = rec[4], Bv = rec[6], Cu = rec[7], Cv = rec[8];
            const float smax = fmaxf(fabsf(rec[9]), fmaxf(fabsf(rec[11]), fabsf(rec[13])));
            const float U = fmaxf(fabsf(Au), fabsf(Cu)) + 1.0f;
            const float V = fmaxf(fabsf(Av), fmaxf(fabsf(Bv), fabsf(Cv))) + 1.0f;
            const float x = 8.0f * (3.0f * smax * U + 3.0f * V) * (1.0f / 16777216.0f);
            const int margin = 3 + (int)fminf(x * 1.001f, 1.0e6f);
            const int vmin = (int)fminf(Av, fminf(Bv, Cv)), vmax = (int)fmaxf(Av, fmaxf(Bv, Cv));
            int4 bx = make_int4(max((int)Au, 0), min((int)Cu, P.W), max(vmin - margin, 0), min(vmax + margin + 1, P.H));
            if (bx.x >= bx.y || bx.z >= bx.w) bx = make_int4(0, 0, 0, 0);
            ((int4*)P.tbox[img])[ti] = bx;
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
// listed triangles, 16 chunks per round (their boxes are tested against the
// tile, a lane per triangle, when the mask alone does not place a triangle
// in the tile).  Wave w then takes the gathered triangles w, w + 4, ...: lane
// i loads the record of the wave's i-th triangle, and the wave walks them one
// at a time (readlane), a lane per tile column u in the triangle's half-open
// column range [A_u, C_u): the column's row span between the long edge AC
// and AB (u < B_u) or BC, as the reference computes it, clipped to the tile.
// Each pixel of the span takes the LDS atomic max of the triangle index:
// where spans overlap (the reference races there) the highest index wins,
// the result of painting the triangles in index order.  Last, every pixel of
// the tile is written once: the plane prior (int)(a*u + b*v + c) of its
// triangle, or -1 where no span covers it.  No global atomics.
#define ELAS_RT_HITS 1024  // 16 chunks of 64 triangles per gather round
__global__ __launch_bounds__(256) void k_rtile(ElasLaunch L) {
    __shared__ int32_t ids[ELAS_RT_H * ELAS_RT_W];
    __shared__ int32_t hits[ELAS_RT_HITS];
    __shared__ unsigned long long nzm[256];
    __shared__ int32_t nzc[256];
    __shared__ int32_t nh;
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
        // the count and the list's first 64 entries in one round trip
        const int nlist = P.scnt[img][st];
        const int4 e0 = tid < 64 && tid < P.scap ? slist[tid] : make_int4(0, 0, 0, 0);
        const int ntri = P.ntri[img], nch = (ntri + 63) >> 6;
        const bool over = nlist > P.scap;        // incomplete list: every chunk, every triangle
        const bool boxed = sh != 0 || over;      // a mask bit does not imply the tile
        const int nent = over ? nch : nlist;
        const int4* box = (const int4*)P.tbox[img];
        const float* trirec = P.trirec[img];
        const int u = u0 + lane;
        const float fu = (float)u;
        for (int cb = 0; cb < nent; cb += 256) {
            // up to 256 chunks per pass
            const int j = cb + tid;
            if (j < nent) {
                if (over) {
                    nzc[tid] = j;
                    nzm[tid] = ntri - 64 * j >= 64 ? ~0ull : (1ull << (ntri - 64 * j)) - 1ull;
                } else {
                    const int4 e = j < 64 ? e0 : slist[j];
                    nzc[tid] = e.x;
                    nzm[tid] = (unsigned long long)(uint32_t)e.y | ((unsigned long long)(uint32_t)e.z << 32);
                }
            }
            __syncthreads();
            const int nz = min(256, nent - cb);
            for (int r = 0; r < nz; r += 16) {
                if (tid == 0) nh = 0;
                __syncthreads();
                // gather: wave w takes chunks r + w + 4q, a lane per triangle
                int cq[4];
                bool in[4];
                int4 bq[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int jq = r + wave + 4 * q;
                    cq[q] = jq < nz ? nzc[jq] : 0;
                    in[q] = jq < nz && ((nzm[jq] >> lane) & 1ull);
                    bq[q] = in[q] && boxed ? box[cq[q] * 64 + lane] : make_int4(u0, u0 + 1, v0, v0 + 1);
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const bool hit = in[q] && bq[q].x < u0 + ELAS_RT_W && bq[q].y > u0 &&
                                     bq[q].z < v0 + ELAS_RT_H && bq[q].w > v0;
                    const unsigned long long bal = __ballot(hit);
                    if (bal) {
                        int base = 0;
                        if (lane == 0) base = atomicAdd(&nh, __popcll(bal));
                        base = __builtin_amdgcn_readfirstlane(base);
                        const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0));
                        if (hit) hits[base + below] = cq[q] * 64 + lane;
                    }
                }
                __syncthreads();
                const int n = nh;
                const int cnt = n > wave ? (n - wave + 3) >> 2 : 0;
                for (int g = 0; g < cnt; g += 64) {
                    int tl = 0;
                    float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0, r2 = r0, r3 = r0;
                    if (g + lane < cnt) {
                        tl = hits[wave + 4 * (g + lane)];
                        const float4* rp = (const float4*)(trirec + (size_t)tl * ELAS_TRIREC);
                        r0 = rp[0]; r1 = rp[1]; r2 = rp[2]; r3 = rp[3];
                    }
                    const int gn = min(64, cnt - g);
                    for (int h = 0; h < gn; h++) {
                        const int Au = (int)__int_as_float(__builtin_amdgcn_readlane(__float_as_int(r0.w), h));
                        const int Bu = (int)__int_as_float(__builtin_amdgcn_readlane(__float_as_int(r1.y), h));
                        const int Cu = (int)__int_as_float(__builtin_amdgcn_readlane(__float_as_int(r1.w), h));
                        if (u >= max(Au, 0) && u < min(Cu, W)) {
                            const int ti = __builtin_amdgcn_readlane(tl, h);
                            const float ABa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r2.y), h));
                            const float ABb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r2.z), h));
                            const float ACa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r2.w), h));
                            const float ACb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r3.x), h));
                            const float BCa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r3.y), h));
                            const float BCb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r3.z), h));
                            const bool first = u < Bu;
                            const float ea = first ? ABa : BCa, eb = first ? ABb : BCb;
                            const int v1 = (int)elas_f2u_sat(__fadd_rn(__fmul_rn(ACa, fu), ACb));
                            const int v2 = (int)elas_f2u_sat(__fadd_rn(__fmul_rn(ea, fu), eb));
                            const int vlo = max(max(min(v1, v2), 0), v0);
                            const int vhi = min(min(max(v1, v2), H), v0 + ELAS_RT_H);
                            for (int v = vlo; v < vhi; v++) atomicMax(&ids[(v - v0) * ELAS_RT_W + lane], ti);
                        }
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
        int32_t* prior = P.prior[img];
        for (int i = tid; i < ELAS_RT_H * ELAS_RT_W; i += 256) {
            const int uu = u0 + (i % ELAS_RT_W), v = v0 + i / ELAS_RT_W;
            if (uu < W && v < H) {
                const int id = ids[i];
                const float* rec = trirec + (size_t)(id >= 0 ? id : 0) * ELAS_TRIREC;
                const int d = elas_f2i_sat(__fadd_rn(__fadd_rn(__fmul_rn(rec[0], (float)uu), __fmul_rn(rec[1], (float)v)), rec[2]));
                prior[(size_t)v * W + uu] = id >= 0 ? d : -1;
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
// is a (row, column) of a cell, so the cell's candidate set is wave-uniform
// and is walked on the scalar unit as the set bits of its bitmask.
//
// LDS holds two rings of R descriptor rows:
//   ring 0: desc0 over [u_lo, u_hi + dmax): the left pixels' own blocks and
//           the right pixels' matches at u + d;
//   ring 1: desc1 over [u_lo - dmax, u_hi): the right pixels' own blocks and
//           the left pixels' matches at u - d.
// A ring row has RW slots, the smallest power of two >= tw + dmax.  Column c
// of step s sits at logical slot (ph + c - base) mod RW, with
// ph = (s - s0)*tw mod RW and base = u_lo or u_lo - dmax, so the tw columns a
// step brings in take the slots of columns that have left, and each
// descriptor row of a chunk is read from HBM once.  With RW a power of two a
// slot address is one subtraction and one and-or.  Row l of the rings is
// rotated by 4 l slots so that a wave's rows start on different banks.  The
// next step's columns, priors and candidate masks are loaded into registers
// while the current step matches and are written after it.
//
// Reference order and tie rule: grid candidates outside the plane window
// [dp-r, dp+r] ascending, then the window ascending with the prior P added,
// strict '<'.  This equals the minimum of the key
//   ((cost + bias) << 11) | (phase << 10) | d,   phase 0 = grid, 1 = window.
// When every P(k) < 0, a grid candidate inside the window is dominated by its
// own window entry (same d, lower cost), so phase 0 need not exclude the
// window; otherwise it does.

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

// CPW = P.dcpw: cells per wave per step (a launch holds the pairs of one CPW).
// Every global load and store is a buffer access issued without a branch
// (lanes with nothing to move take an out-of-range offset), so the waits
// below are for exactly the loads they need.
template <int CPW, bool PNEG>
__global__ __launch_bounds__(ELAS_DENSE_THREADS, CPW == 1 ? 6 : 4) void k_dense(ElasLaunch L, ElasKParams kp) {
    extern __shared__ uint4 s_ring[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar registers
    // a candidate outside the image (or of an excluded kind) starts its SAD
    // from KEY_BAD: its key exceeds every valid key and stays below 2^32
    constexpr uint32_t KEY_BAD = 0x1ff000u;
    // ring reads by LDS byte address: the dynamic LDS (the rings) starts
    // right after the kernel's static LDS, of which it has none, so at 0
    typedef const __attribute__((address_space(3))) elas_u32x4 lds_u4;
    constexpr uint32_t lds0 = 0u;
    for (int t0 = blockIdx.x; t0 < L.total; t0 += gridDim.x) {
        const int t = elas_xcd_order(t0, L.total, gridDim.x);
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int W = P.W, H = P.H, gs = P.gsize, gw = P.gw, dmax = P.dmax, words = P.words;
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
        // ring rows of RW = 2^k >= tw + dmax slots; line l (ring 0 rows, then
        // ring 1 rows) is rotated by 4 l slots so that a wave's rows start
        // on different banks: logical slot q of line l sits at
        // l*RW + ((q + 4 l) mod RW)
        const int RW = (int)elas_pow2_at_least((uint32_t)(tw + dmax));
        const uint32_t M16 = 16u * (uint32_t)RW - 1u;
        auto phys = [&](int line, int q) { return line * RW + ((q + 4 * line) & (RW - 1)); };
        const uint32_t npx = (uint32_t)W * (uint32_t)H;  // <= 2^27
        const __amdgpu_buffer_rsrc_t rd0 = elas_rsrc(P.desc[0], 16u * npx);
        const __amdgpu_buffer_rsrc_t rd1 = elas_rsrc(P.desc[1], 16u * npx);
        // descriptor of image k, band row rr, column c (0 off the image); k is
        // a constant at every call, so the buffer resource stays scalar
        auto dload = [&](int k, int rr, int c, bool ok) {
            const uint32_t off = ok && c >= 0 && c < W ? ((uint32_t)(v0 + rr) * (uint32_t)W + (uint32_t)c) * 16u : ELAS_OOR;
            return elas_ld16_ranged(k ? rd1 : rd0, off);
        };
        const int row = lane / gs, col = lane - row * gs;
        const int lr = row < rows ? row : 0;
        const int v = v0 + row;
        const bool lane_on = lane < R * gs && row < rows;
        const uint32_t vrow4 = 4u * (uint32_t)v * (uint32_t)W;  // byte offset of this lane's row in a float/int map
        const __amdgpu_buffer_rsrc_t rp[2] = {elas_rsrc(P.prior[0], 4u * npx), elas_rsrc(P.prior[1], 4u * npx)};
        const uint32_t mbytes = 4u * (uint32_t)(P.gh * gw * words);
        const __amdgpu_buffer_rsrc_t rm[2] = {elas_rsrc(P.gmask[0], mbytes), elas_rsrc(P.gmask[1], mbytes)};
        const __amdgpu_buffer_rsrc_t ro[2] = {elas_rsrc(P.draw[0], 4u * npx), elas_rsrc(P.draw[1], 4u * npx)};
        // priors of this lane's pixel in each (image, cell) of step s (0 off
        // the matching domain), and the cells' candidate masks: lane 32*ci + w
        // holds word w of cell ci
        auto load_px = [&](int s, int (&pr)[2][CPW], uint32_t (&mk)[2]) {
            #pragma unroll
            for (int img = 0; img < 2; img++) {
                const int ci = lane >> 5, w = lane & 31;
                const int gx = s * tcells + wave + ci * ELAS_DENSE_WAVES;
                const bool okm = ci < CPW && w < words && gx < gw && s < s1;
                mk[img] = elas_ld4_ranged(rm[img], okm ? ((uint32_t)(gy * gw + gx) * (uint32_t)words + (uint32_t)w) * 4u
                                                       : ELAS_OOR);
                #pragma unroll
                for (int c = 0; c < CPW; c++) {
                    const int gxc = s * tcells + wave + c * ELAS_DENSE_WAVES;
                    const int u = gxc * gs + col;
                    const bool ok = s < s1 && gxc < gw && lane_on && u >= 4 && u <= W - 4 && v >= 4 && v <= H - 4;
                    pr[img][c] = (int)elas_ld4_ranged(rp[img], ok ? vrow4 + 4u * (uint32_t)u : ELAS_OOR);
                }
            }
        };
        // this step's priors and masks are in flight with the ring loads below
        int pr_cur[2][CPW];
        uint32_t mk_cur[2];
        load_px(s0, pr_cur, mk_cur);
        // prologue: step s0's windows (slot = column - base), both rings at
        // the same (row, slot), every load of a batch issued before its stores
        {
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
                        s_ring[phys(li[e] >> 16, li[e] & 0xffff)] = q[0][e];
                        s_ring[phys(rows + (li[e] >> 16), li[e] & 0xffff)] = q[1][e];
                    }
            }
        }
        // this thread's share of a step's new columns, the same (row, column)
        // in both rings (rows x tw <= 64 * 8 * CPW per ring):
        // (row << 16) | column, or -1
        int pf_at[CPW];
        {
            const int n = rows * tw;
            #pragma unroll
            for (int e = 0; e < CPW; e++) {
                const int i = tid + e * ELAS_DENSE_THREADS;
                pf_at[e] = i < n ? ((i / tw) << 16) | (i - (i / tw) * tw) : -1;
            }
        }
        elas_lds_barrier();
        int ph = 0;
        for (int s = s0; s < s1; s++) {
            // next step's new columns (ring 0: [u + dmax, u + dmax + tw),
            // ring 1: [u, u + tw) with u = (s + 1) tw), priors and masks; past
            // the chunk's last step these loads read nothing
            const bool more = s + 1 < s1;
            const int u_nx = (s + 1) * tw;
            uint4 pf[2][CPW];
            #pragma unroll
            for (int e = 0; e < CPW; e++) {
                const int rr = pf_at[e] >> 16, j = pf_at[e] & 0xffff;
                pf[0][e] = dload(0, rr, u_nx + dmax + j, more && pf_at[e] >= 0);
                pf[1][e] = dload(1, rr, u_nx + j, more && pf_at[e] >= 0);
            }
            int pr_nx[2][CPW];
            uint32_t mk_nx[2];
            load_px(s + 1, pr_nx, mk_nx);
            #pragma unroll
            for (int img = 0; img < 2; img++) {
                const int sgn = img ? 1 : -1;
                // own blocks in ring img, matches in ring 1 - img: byte address of
                // logical slot q of line l = ((16 (q + 4 l)) & M16) | 16 RW l
                const int own_line = img * rows + lr, oth_line = (1 - img) * rows + lr;
                const uint32_t own_base = 16u * (uint32_t)(RW * own_line), oth_base = 16u * (uint32_t)(RW * oth_line);
                auto ld = [&](uint32_t a) {
                    const elas_u32x4 x = *(lds_u4*)(uintptr_t)(lds0 + a);
                    return make_uint4(x.x, x.y, x.z, x.w);
                };
                #pragma unroll
                for (int ci = 0; ci < CPW; ci++) {
                    const int gx = s * tcells + wave + ci * ELAS_DENSE_WAVES;
                    const int u = gx * gs + col;
                    const bool active = lane_on && gx < gw && u < W;
                    float res = -1.0f;
                    if (gx < gw) {  // wave-uniform
                        const int xo = (gx - s * tcells) * gs + col;  // u - u_lo
                        const int dp = pr_cur[img][ci];
                        const uint4 b1 = ld((16u * (uint32_t)(ph + 4 * own_line + xo + (img ? dmax : 0)) & M16) | own_base);
                        bool valid = active && u >= 4 && u <= W - 4 && v >= 4 && v <= H - 4 && dp > 0;
                        if (kp.match_texture > 0) {
                            const uint32_t tex =
                                elas_sad16(b1, make_uint4(0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u), 0u);
                            if ((int)tex < kp.match_texture) valid = false;
                        }
                        // u_warp = u -/+ d must lie in [3, W-4] (KerComputeDisparity.cu:259,267): d <= dlim
                        const int dlim = img ? W - 4 - u : u - 3;
                        // the match at disparity d: logical slot bm + sgn*d
                        const uint32_t bm = 16u * (uint32_t)(ph + 4 * oth_line + xo + (img ? 0 : dmax));
                        auto maddr = [&](int d) { return ((bm + (uint32_t)(sgn * 16 * d)) & M16) | oth_base; };
                        // every column of the cell takes d <= dlim_min
                        const int dlim_min = img ? W - 4 - (gx * gs + gs - 1) : gx * gs - 3;
                        const uint32_t bias = (uint32_t)kp.key_bias;
                        const int dpmin = max(dp - kp.plane_radius, 0), dpmax = min(dp + kp.plane_radius, dmax);
                        uint32_t kmin = 0xffffffffu;
                        // grid candidates (wave-uniform d): key = (sad + bias) << 11 | d
                        auto grid_acc = [&](int d) {
                            bool ok = d <= dlim;
                            if (!PNEG) ok = ok && (d < dpmin || d > dpmax);
                            return ok ? bias : KEY_BAD;
                        };
                        const uint32_t mkv = mk_cur[img];
                        // candidates in ascending d across the cell's non-empty mask
                        // words (one ballot finds them), four at a time so that their
                        // LDS reads overlap, then the last one to three singly
                        uint32_t wleft = (uint32_t)(__ballot(mkv != 0u) >> (32 * ci));
                        if (words < 32) wleft &= (1u << words) - 1u;
                        int left = 0;
                        for (uint32_t m = wleft; m; m &= m - 1u)
                            left += __builtin_popcount(__builtin_amdgcn_readlane(mkv, 32 * ci + __builtin_ctz(m)));
                        uint32_t bits = 0u;
                        int wbase = 0;
                        auto next = [&]() {
                            while (!bits) {
                                const int w = __builtin_ctz(wleft);
                                wleft &= wleft - 1u;
                                bits = __builtin_amdgcn_readlane(mkv, 32 * ci + w);
                                wbase = 32 * w;
                            }
                            const int d = wbase + __builtin_ctz(bits);
                            bits &= bits - 1u;
                            return d;
                        };
                        for (; left >= 4; left -= 4) {
                            int dq[4];
                            uint4 b2[4];
                            #pragma unroll
                            for (int q = 0; q < 4; q++) {
                                dq[q] = next();
                                b2[q] = ld(maddr(dq[q]));
                            }
                            if (PNEG && dq[3] <= dlim_min) {  // wave-uniform: no per-lane test
                                #pragma unroll
                                for (int q = 0; q < 4; q++)
                                    kmin = min(kmin, (elas_sad16(b1, b2[q], bias) << 11) | (uint32_t)dq[q]);
                            } else {
                                #pragma unroll
                                for (int q = 0; q < 4; q++)
                                    kmin = min(kmin, (elas_sad16(b1, b2[q], grid_acc(dq[q])) << 11) | (uint32_t)dq[q]);
                            }
                        }
                        for (; left > 0; left--) {
                            const int d = next();
                            kmin = min(kmin, (elas_sad16(b1, ld(maddr(d)), grid_acc(d)) << 11) | (uint32_t)d);
                        }
                        // plane window d = dp - r + k, k = 0..2r, with the prior P(|k - r|)
                        // added: key = (sad + P + bias) << 11 | 1024 | d
                        const int pr = kp.plane_radius, nwin = 2 * pr + 1;
                        const int wlim = min(dmax, dlim);
                        const int d0 = dp - pr;
                        auto win_key = [&](int k, const uint4 b2) {
                            const int d = d0 + k;
                            const uint32_t acc =
                                (uint32_t)d <= (uint32_t)wlim ? (uint32_t)((int)bias + kp.P[abs(k - pr)]) : KEY_BAD;
                            return (elas_sad16(b1, b2, acc) << 11) | (1024u | ((uint32_t)d & 1023u));
                        };
                        // a window entry off [0, wlim] starts from KEY_BAD, so its
                        // (wrapped, in-bounds) read need not be clamped
                        const uint32_t aw = bm + (uint32_t)(sgn * 16 * d0);
                        auto win_lds = [&](int k) { return ld(((aw + (uint32_t)(sgn * 16 * k)) & M16) | oth_base); };
                        int k0 = 0;
                        if (nwin == 5) {
                            // the effective parameters: one batch of five, the priors
                            // P(2), P(1), P(0), P(1), P(2) indexed by constants (a
                            // runtime index into the kernel arguments is a scalar
                            // memory load per entry)
                            const int pw[5] = {kp.P[2], kp.P[1], kp.P[0], kp.P[1], kp.P[2]};
                            uint4 b2[5];
                            #pragma unroll
                            for (int q = 0; q < 5; q++) b2[q] = win_lds(q);
                            // when every valid lane's window lies in [0, wlim] (the usual
                            // case) the prior is added without a per-lane test; other
                            // lanes' results are discarded
                            const bool inside = d0 >= 0 && d0 + 4 <= wlim;
                            if (__ballot(valid && !inside) == 0) {
                                const uint32_t kd = 1024u | (uint32_t)d0;
                                #pragma unroll
                                for (int q = 0; q < 5; q++)
                                    kmin = min(kmin, (elas_sad16(b1, b2[q], (uint32_t)((int)bias + pw[q])) << 11) |
                                                         ((kd + (uint32_t)q) & 2047u));
                            } else {
                                #pragma unroll
                                for (int q = 0; q < 5; q++) {
                                    const int d = d0 + q;
                                    const uint32_t acc =
                                        (uint32_t)d <= (uint32_t)wlim ? (uint32_t)((int)bias + pw[q]) : KEY_BAD;
                                    kmin = min(kmin, (elas_sad16(b1, b2[q], acc) << 11) | (1024u | ((uint32_t)d & 1023u)));
                                }
                            }
                            k0 = 5;
                        }
                        for (; k0 + 4 <= nwin; k0 += 4) {
                            uint4 b2[4];
                            #pragma unroll
                            for (int q = 0; q < 4; q++) b2[q] = win_lds(k0 + q);
                            #pragma unroll
                            for (int q = 0; q < 4; q++) kmin = min(kmin, win_key(k0 + q, b2[q]));
                        }
                        for (; k0 < nwin; k0++) kmin = min(kmin, win_key(k0, win_lds(k0)));
                        if (valid && kmin < (KEY_BAD << 11)) res = (float)(kmin & 1023u);
                    }
                    elas_st4_ranged(ro[img], active ? vrow4 + 4u * (uint32_t)u : ELAS_OOR, __float_as_uint(res));
                }
            }
            const int phn = (ph + tw) & (RW - 1);
            if (more) {
                elas_lds_barrier();  // every wave is done with step s's slots
                #pragma unroll
                for (int e = 0; e < CPW; e++)
                    if (pf_at[e] >= 0) {
                        const int rr = pf_at[e] >> 16, q = phn + dmax + (pf_at[e] & 0xffff);
                        s_ring[phys(rr, q)] = pf[0][e];
                        s_ring[phys(rows + rr, q)] = pf[1][e];
                    }
                elas_lds_barrier();
            }
            #pragma unroll
            for (int img = 0; img < 2; img++) {
                mk_cur[img] = mk_nx[img];
                #pragma unroll
                for (int c = 0; c < CPW; c++) pr_cur[img][c] = pr_nx[img][c];
            }
            ph = phn;
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
    if (cpw == 2) {
        if (kp.all_p_negative) hipLaunchKernelGGL((k_dense<2, true>), dim3(grid), b, lds, s, L, kp);
        else hipLaunchKernelGGL((k_dense<2, false>), dim3(grid), b, lds, s, L, kp);
    } else {
        if (kp.all_p_negative) hipLaunchKernelGGL((k_dense<1, true>), dim3(grid), b, lds, s, L, kp);
        else hipLaunchKernelGGL((k_dense<1, false>), dim3(grid), b, lds, s, L, kp);
    }
}
int elas_dense_allow_lds(size_t bytes) {
    const void* fn[4] = {reinterpret_cast<const void*>(&k_dense<1, true>), reinterpret_cast<const void*>(&k_dense<1, false>),
                         reinterpret_cast<const void*>(&k_dense<2, true>), reinterpret_cast<const void*>(&k_dense<2, false>)};
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
