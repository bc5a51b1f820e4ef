// k_post.hip -- post-processing filters on the final disparity maps (gfx950).
//
// The reference declares these filters but never implements them
// (speckle_sim_threshold / speckle_size / ipol_gap_width / filter_median /
// filter_adaptive_mean, src/Elas.h:31-35, src/Kernels/KernelsCommon.h:19-21;
// its pipeline ends at the LR check, src/Elas.cu:107-119).  Their definitions
// are this project's own, written from the ELAS paper (Geiger, Roser and
// Urtasun, ACCV 2010); values < 0 are invalid throughout.
//
//   speckle removal  4-connected segments of valid pixels, neighbours joined
//                    when |d_a - d_b| <= speckle_sim; segments with fewer than
//                    speckle_size pixels become -1.  Union-find: each 64x16
//                    tile joins its pixels in LDS, tile borders are joined in
//                    HBM with atomicMin on roots, then one pass counts segment
//                    sizes (wave-aggregated atomics) and one applies them.  The
//                    partition, and so the result, does not depend on the
//                    order in which unions happen.
//   gap interpolation  rows, then columns of the row-filled map: a run of
//                    invalid pixels with a valid pixel on both sides and length
//                    <= ipol_gap_width gets the mean of the two when they
//                    differ by < 3, else the smaller.  One workgroup per line:
//                    the line sits in LDS, nearest valid pixels left/right of
//                    every thread's chunk come from max/min scans.
//   adaptive mean    rows, then columns: weighted mean over +-3 valid pixels
//                    with weight max(0, 4 - |d_k - d_c|).
//   median           rows, then columns: median of the valid values in +-3
//                    (the lower middle one for even counts).
// The separable filters run both passes in one kernel: a 64x16 output tile
// stages a (16+6)-row by (64+6)-column window in LDS, computes the row pass
// for all 22 rows, then the column pass; the map is read and written once.
// Summation order and the float operations match the CPU definitions, so the
// results are bit-identical to them.
#include <hip/hip_runtime.h>
#include "elas_common.h"
#include "elas_device.h"

#define PT_W ELAS_POST_TW
#define PT_H ELAS_POST_TH
#define PT_R 3  // window radius of the separable filters

namespace {

// tile t of a post-tile stage -> pair, image, tile origin
struct PostTile {
    int p, img, u0, v0;
};
__device__ __forceinline__ PostTile post_tile(const ElasLaunch& L, int t) {
    PostTile r;
    r.p = elas_find_pair(L.prefix, L.npairs, t);
    const auto& P = L.pairs[r.p];
    const int tx = (P.W + PT_W - 1) / PT_W, ty = (P.H + PT_H - 1) / PT_H;
    int local = t - L.prefix[r.p];
    r.img = local / (tx * ty);
    local -= r.img * tx * ty;
    r.u0 = (local % tx) * PT_W;
    r.v0 = (local / tx) * PT_H;
    return r;
}

__device__ __forceinline__ int32_t ld_relaxed(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x, halving the path on the way (atomicMin keeps every link pointing
// at an ancestor, whatever other threads do meanwhile).
__device__ int uf_find_global(int32_t* lab, int x) {
    for (;;) {
        const int p = ld_relaxed(lab + x);
        if (p == x) return x;
        const int g = ld_relaxed(lab + p);
        if (g == p) return p;
        atomicMin(lab + x, g);
        x = g;
    }
}

// Joins the sets of a and b: link the larger root under the smaller with
// atomicMin; a root that was relinked meanwhile is retried from its new parent.
__device__ void uf_union_global(int32_t* lab, int a, int b) {
    for (;;) {
        a = uf_find_global(lab, a);
        b = uf_find_global(lab, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(lab + a, b);
        if (old == a) return;
        a = old;
    }
}

__device__ __forceinline__ int uf_find_lds(int32_t* lab, int x) {
    int y = __hip_atomic_load(lab + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (y != x) {
        x = y;
        y = __hip_atomic_load(lab + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return x;
}

__device__ void uf_union_lds(int32_t* lab, int a, int b) {
    for (;;) {
        a = uf_find_lds(lab, a);
        b = uf_find_lds(lab, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(lab + a, b);
        if (old == a) return;
        a = old;
    }
}

}  // namespace

// Tile-local union-find; writes every pixel's tile root as a global index
// (-1 for invalid pixels) and clears the size array.
__global__ __launch_bounds__(256) void k_speckle_local(ElasLaunch L, ElasKParams kp) {
    __shared__ float s_d[PT_H][PT_W];
    __shared__ int32_t s_lab[PT_H * PT_W];
    const int tid = threadIdx.x;
    const float sim = kp.speckle_sim;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const PostTile T = post_tile(L, t);
        if (T.img >= kp.post_images) continue;
        const auto& P = L.pairs[T.p];
        const int W = P.W, H = P.H;
        const float* D = P.dout[T.img];
        float q[PT_H * PT_W / 256];
        #pragma unroll
        for (int k = 0; k < PT_H * PT_W / 256; k++) {
            const int i = tid + 256 * k, r = i / PT_W, c = i % PT_W;
            const int u = T.u0 + c, v = T.v0 + r;
            q[k] = (u < W && v < H) ? elas_ld(D + (size_t)v * W + u) : -1.0f;
        }
        #pragma unroll
        for (int k = 0; k < PT_H * PT_W / 256; k++) {
            const int i = tid + 256 * k;
            s_d[i / PT_W][i % PT_W] = q[k];
            s_lab[i] = q[k] >= 0.0f ? i : -1;
        }
        elas_lds_barrier();
        #pragma unroll
        for (int k = 0; k < PT_H * PT_W / 256; k++) {
            const int i = tid + 256 * k, r = i / PT_W, c = i % PT_W;
            const float d = q[k];
            if (d < 0.0f) continue;
            if (c + 1 < PT_W && s_d[r][c + 1] >= 0.0f && fabsf(d - s_d[r][c + 1]) <= sim)
                uf_union_lds(s_lab, i, i + 1);
            if (r + 1 < PT_H && s_d[r + 1][c] >= 0.0f && fabsf(d - s_d[r + 1][c]) <= sim)
                uf_union_lds(s_lab, i, i + PT_W);
        }
        elas_lds_barrier();
        int32_t* lab = P.pbuf[T.img];
        int32_t* sz = lab + (size_t)W * H;
        #pragma unroll
        for (int k = 0; k < PT_H * PT_W / 256; k++) {
            const int i = tid + 256 * k, r = i / PT_W, c = i % PT_W;
            const int u = T.u0 + c, v = T.v0 + r;
            if (u >= W || v >= H) continue;
            int g = -1;
            if (q[k] >= 0.0f) {
                const int root = uf_find_lds(s_lab, i);
                g = (T.v0 + root / PT_W) * W + T.u0 + root % PT_W;
            }
            elas_st(lab + (size_t)v * W + u, g);
            elas_st(sz + (size_t)v * W + u, 0);
        }
        elas_lds_barrier();
    }
}

// Joins segments across tile borders (right column and bottom row of a tile).
__global__ __launch_bounds__(256) void k_speckle_merge(ElasLaunch L, ElasKParams kp) {
    const int tid = threadIdx.x;
    const float sim = kp.speckle_sim;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const PostTile T = post_tile(L, t);
        if (T.img >= kp.post_images) continue;
        const auto& P = L.pairs[T.p];
        const int W = P.W, H = P.H;
        const float* D = P.dout[T.img];
        int32_t* lab = P.pbuf[T.img];
        // threads 0..PT_H-1: right border rows; PT_H..PT_H+PT_W-1: bottom border columns
        int u = -1, v = -1, un = -1, vn = -1;
        if (tid < PT_H) {
            u = T.u0 + PT_W - 1; v = T.v0 + tid; un = u + 1; vn = v;
        } else if (tid < PT_H + PT_W) {
            u = T.u0 + tid - PT_H; v = T.v0 + PT_H - 1; un = u; vn = v + 1;
        }
        if (u < 0 || u >= W || v >= H || un >= W || vn >= H) continue;
        const float a = D[(size_t)v * W + u], b = D[(size_t)vn * W + un];
        if (a >= 0.0f && b >= 0.0f && fabsf(a - b) <= sim)
            uf_union_global(lab, v * W + u, vn * W + un);
    }
}

// Final roots and segment sizes.
__global__ __launch_bounds__(256) void k_speckle_count(ElasLaunch L, ElasKParams kp) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int n = P.W * P.H;
        const int i = (t - L.prefix[p]) * 256 + threadIdx.x;
        for (int k = 0; k < kp.post_images; k++) {
            int32_t* lab = P.pbuf[k];
            int32_t* sz = lab + n;
            int r = -1;
            if (i < n && ld_relaxed(lab + i) >= 0) {
                r = uf_find_global(lab, i);
                // the root is the smallest index of its set: readers see an ancestor either way
                __hip_atomic_store(lab + i, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            // one atomic per distinct root in the wave
            bool todo = r >= 0;
            while (__any(todo)) {
                const int r0 = __shfl(r, __ffsll((unsigned long long)__ballot(todo)) - 1);
                const unsigned long long m = __ballot(todo && r == r0);
                if (todo && r == r0) {
                    if ((int)(__lane_id()) == __ffsll(m) - 1) atomicAdd(sz + r0, (int)__popcll(m));
                    todo = false;
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_speckle_apply(ElasLaunch L, ElasKParams kp) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int n = P.W * P.H;
        const int i = (t - L.prefix[p]) * 256 + threadIdx.x;
        if (i >= n) continue;
        for (int k = 0; k < kp.post_images; k++) {
            const int32_t* lab = P.pbuf[k];
            const int r = lab[i];
            if (r >= 0 && lab[n + r] < kp.speckle_size) P.dout[k][i] = -1.0f;
        }
    }
}

// Gap interpolation along rows (COLS = false) or columns (COLS = true); one
// workgroup per line, the line in LDS (dynamic, line length floats).
template <bool COLS>
__global__ __launch_bounds__(256) void k_gap(ElasLaunch L, ElasKParams kp) {
    extern __shared__ float s_line[];
    __shared__ int32_t s_last[256], s_first[256];
    const int tid = threadIdx.x;
    const int gap = kp.ipol_gap;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int W = P.W, H = P.H;
        const int nl = COLS ? W : H, len = COLS ? H : W;
        const int local = t - L.prefix[p];
        const int img = local / nl, l = local - img * nl;
        if (img >= kp.post_images) continue;
        float* D = P.dout[img];
        const size_t es = COLS ? (size_t)W : 1, base = COLS ? (size_t)l : (size_t)l * W;
        for (int i = tid; i < len; i += 256) s_line[i] = D[base + i * es];
        elas_lds_barrier();
        const int C = (len + 255) / 256;
        const int c0 = min(len, tid * C), c1 = min(len, c0 + C);
        int lastv = -1, firstv = 0x7fffffff;
        for (int i = c0; i < c1; i++)
            if (s_line[i] >= 0.0f) {
                lastv = i;
                firstv = min(firstv, i);
            }
        s_last[tid] = lastv;
        s_first[tid] = firstv;
        elas_lds_barrier();
        // inclusive prefix max of s_last, inclusive suffix min of s_first
        for (int off = 1; off < 256; off <<= 1) {
            const int a = tid >= off ? s_last[tid - off] : -1;
            const int b = tid + off < 256 ? s_first[tid + off] : 0x7fffffff;
            elas_lds_barrier();
            s_last[tid] = max(s_last[tid], a);
            s_first[tid] = min(s_first[tid], b);
            elas_lds_barrier();
        }
        int Lv = tid > 0 ? s_last[tid - 1] : -1;
        const int Rin = tid < 255 ? s_first[tid + 1] : 0x7fffffff;
        int q = c0;
        while (q < c1) {
            if (s_line[q] >= 0.0f) {
                Lv = q++;
                continue;
            }
            int e = q;
            while (e < c1 && s_line[e] < 0.0f) e++;
            const int j = e < c1 ? e : Rin;  // first valid pixel after the run
            if (Lv >= 0 && j < len && j - (Lv + 1) <= gap) {
                const float d1 = s_line[Lv], d2 = s_line[j];
                const float f = fabsf(d1 - d2) < 3.0f ? (d1 + d2) * 0.5f : fminf(d1, d2);
                for (int k = q; k < e; k++) s_line[k] = f;
            }
            q = e;
        }
        elas_lds_barrier();
        for (int i = tid; i < len; i += 256) D[base + i * es] = s_line[i];
        elas_lds_barrier();
    }
}

namespace {

// weighted mean over the +-3 window w[0..6] around w[3]; < 0 = skip
__device__ __forceinline__ float adaptive_1d(const float* w) {
    const float dc = w[PT_R];
    if (dc < 0.0f) return dc;
    float ws = 0.0f, acc = 0.0f;
    #pragma unroll
    for (int k = 0; k < 2 * PT_R + 1; k++) {
        const float dk = w[k];
        if (dk < 0.0f) continue;
        const float wt = fmaxf(0.0f, 4.0f - fabsf(dk - dc));
        ws = ws + wt;
        acc = acc + wt * dk;
    }
    return acc / ws;
}

__device__ __forceinline__ void cswap(float& a, float& b) {
    const float lo = fminf(a, b), hi = fmaxf(a, b);
    a = lo;
    b = hi;
}

// lower median of the valid values of w[0..6] (w[3] valid)
__device__ __forceinline__ float median_1d(const float* w) {
    const float dc = w[PT_R];
    if (dc < 0.0f) return dc;
    float b[7];
    int n = 0;
    #pragma unroll
    for (int k = 0; k < 7; k++) {
        const bool ok = w[k] >= 0.0f;
        b[k] = ok ? w[k] : __builtin_inff();
        n += ok;
    }
    // sorting network for 7 inputs (16 compare-exchanges)
    cswap(b[0], b[6]); cswap(b[2], b[3]); cswap(b[4], b[5]);
    cswap(b[0], b[2]); cswap(b[1], b[4]); cswap(b[3], b[6]);
    cswap(b[0], b[1]); cswap(b[2], b[5]); cswap(b[3], b[4]);
    cswap(b[1], b[2]); cswap(b[4], b[6]);
    cswap(b[2], b[3]); cswap(b[4], b[5]);
    cswap(b[1], b[2]); cswap(b[3], b[4]);
    cswap(b[5], b[6]);
    const int m = (n - 1) >> 1;
    float r = b[0];
    #pragma unroll
    for (int k = 1; k < 4; k++) r = (m == k) ? b[k] : r;
    return r;
}

}  // namespace

// Separable +-3 filter (MEDIAN: median, else adaptive mean), both passes in
// one kernel.  src/dst: 0 = dout, 1 = the scratch map (pbuf + W*H).
template <bool MEDIAN>
__global__ __launch_bounds__(256) void k_sep(ElasLaunch L, ElasKParams kp, int from_tmp) {
    constexpr int RH = PT_H + 2 * PT_R, RW = PT_W + 2 * PT_R;
    __shared__ float s_in[RH][RW + 1];
    __shared__ float s_row[RH][PT_W + 1];
    const int tid = threadIdx.x;
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const PostTile T = post_tile(L, t);
        if (T.img >= kp.post_images) continue;
        const auto& P = L.pairs[T.p];
        const int W = P.W, H = P.H;
        float* mapA = P.dout[T.img];
        float* mapB = reinterpret_cast<float*>(P.pbuf[T.img] + (size_t)W * H);
        const float* src = from_tmp ? mapB : mapA;
        float* dst = from_tmp ? mapA : mapB;
        {
            constexpr int N = RH * RW, K = (N + 255) / 256;
            float q[K];
            #pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = tid + 256 * k, r = i / RW, c = i - r * RW;
                const int u = T.u0 - PT_R + c, v = T.v0 - PT_R + r;
                q[k] = (i < N && u >= 0 && u < W && v >= 0 && v < H) ? elas_ld(src + (size_t)v * W + u) : -1.0f;
            }
            #pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = tid + 256 * k, r = i / RW, c = i - r * RW;
                if (i < N) s_in[r][c] = q[k];
            }
        }
        elas_lds_barrier();
        // row pass over all staged rows (rows outside the image stay absent)
        for (int i = tid; i < RH * PT_W; i += 256) {
            const int r = i / PT_W, c = i - r * PT_W;
            const int v = T.v0 - PT_R + r;
            float w[2 * PT_R + 1];
            #pragma unroll
            for (int k = 0; k < 2 * PT_R + 1; k++) w[k] = s_in[r][c + k];
            float o = -1.0f;
            if (v >= 0 && v < H) o = MEDIAN ? median_1d(w) : adaptive_1d(w);
            s_row[r][c] = o;
        }
        elas_lds_barrier();
        for (int i = tid; i < PT_H * PT_W; i += 256) {
            const int r = i / PT_W, c = i - r * PT_W;
            const int u = T.u0 + c, v = T.v0 + r;
            if (u >= W || v >= H) continue;
            float w[2 * PT_R + 1];
            #pragma unroll
            for (int k = 0; k < 2 * PT_R + 1; k++) w[k] = s_row[r + k][c];
            elas_st(dst + (size_t)v * W + u, MEDIAN ? median_1d(w) : adaptive_1d(w));
        }
        elas_lds_barrier();
    }
}

__global__ __launch_bounds__(256) void k_post_copy(ElasLaunch L, ElasKParams kp) {
    for (int t = blockIdx.x; t < L.total; t += gridDim.x) {
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int n = P.W * P.H;
        const int i = (t - L.prefix[p]) * 256 + threadIdx.x;
        if (i >= n) continue;
        for (int k = 0; k < kp.post_images; k++)
            P.dout[k][i] = reinterpret_cast<const float*>(P.pbuf[k] + n)[i];
    }
}

void elas_launch_speckle(const ElasLaunch& Ltiles, const ElasLaunch& Lpix, const ElasKParams& kp, int grid,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_speckle_local, dim3(min(grid, Ltiles.total)), dim3(256), 0, s, Ltiles, kp);
    hipLaunchKernelGGL(k_speckle_merge, dim3(min(grid, Ltiles.total)), dim3(256), 0, s, Ltiles, kp);
    hipLaunchKernelGGL(k_speckle_count, dim3(min(grid, Lpix.total)), dim3(256), 0, s, Lpix, kp);
    hipLaunchKernelGGL(k_speckle_apply, dim3(min(grid, Lpix.total)), dim3(256), 0, s, Lpix, kp);
}

void elas_launch_gap(const ElasLaunch& Lrows, const ElasLaunch& Lcols, const ElasKParams& kp, int grid,
                     size_t lds_rows, size_t lds_cols, hipStream_t s) {
    hipLaunchKernelGGL(k_gap<false>, dim3(min(grid, Lrows.total)), dim3(256), lds_rows, s, Lrows, kp);
    hipLaunchKernelGGL(k_gap<true>, dim3(min(grid, Lcols.total)), dim3(256), lds_cols, s, Lcols, kp);
}

void elas_launch_sep(const ElasLaunch& Ltiles, const ElasKParams& kp, bool median, int from_tmp, int grid,
                     hipStream_t s) {
    if (median) hipLaunchKernelGGL(k_sep<true>, dim3(min(grid, Ltiles.total)), dim3(256), 0, s, Ltiles, kp, from_tmp);
    else hipLaunchKernelGGL(k_sep<false>, dim3(min(grid, Ltiles.total)), dim3(256), 0, s, Ltiles, kp, from_tmp);
}

void elas_launch_post_copy(const ElasLaunch& Lpix, const ElasKParams& kp, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_post_copy, dim3(min(grid, Lpix.total)), dim3(256), 0, s, Lpix, kp);
}
