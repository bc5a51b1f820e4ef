// k_desc.hip -- fused Sobel + 16-byte descriptor (gfx950).
//
// Restates src/Kernels/KerDescriptor.cu:9-100 (kerComputeSobel and
// kerComputeDescriptor) as one kernel.  du/dv never go to HBM: algorithmic
// traffic is 1 B read + 16 B written per pixel.
//
// du = 128 + clamp(S_u >> 2, -128, 127) is the integer form of the
// reference's (uint8_t)(min(max(-128.f, S/4), 127.f) + 128)
// (KerDescriptor.cu:40-41): S/4 is exact in float and truncating the
// non-negative sum is floor.  Descriptor bytes outside u in [3, W-4],
// v in [3, H-4] are written as 0 (the reference leaves them unwritten).
//
// Work is done on packed bytes, four columns per dword:
//   * a 64x16 output tile stages image rows v0-3 .. v0+18, columns
//     u0-8 .. u0+71, in LDS as dwords (column c of a row = u - u0 + 8);
//   * du and dv for four columns at once: with the vertical (horizontal)
//     weighted sums V = I(r-1) + 2 I(r) + I(r+1) of columns (c, c+2) and
//     (c+1, c+3) held as 16-bit halves of a dword (each <= 1020, so plain
//     32-bit adds never carry between halves), S_u(c) = V(c-1) - V(c+1) and
//     S_v = H(v-1) - H(v+1) are two packed 16-bit subtractions per four
//     columns, then >> 2, clamp and + 128 in packed 16-bit lanes, and one
//     byte permute packs the four results;
//   * each thread writes the descriptors of four adjacent pixels: their 64
//     bytes are gathered with v_perm_b32 (compile-time selectors) from 16
//     dword reads of the du/dv rows around them.
#include <hip/hip_runtime.h>
#include "elas_common.h"
#include "elas_device.h"

#define DT_W ELAS_DESC_TW        // tile width (a multiple of 16); DT_W * DT_H = 1024
#define DT_H (1024 / DT_W)
#define DT_GP (DT_W / 4)         // four-pixel groups per output row
#define DT_NG (DT_GP + 2)        // du/dv groups per row (1 .. DT_GP + 2)
#define DT_DW (DT_GP + 4)        // dwords per staged row (columns u0-8 .. u0+DT_W+7)
#define DT_IR (DT_H + 6)         // image rows v0-3 .. v0+DT_H+2
#define DT_UR (DT_H + 4)         // du rows v0-2 .. v0+DT_H+1
#define DT_VR (DT_H + 2)         // dv rows v0-1 .. v0+DT_H

typedef int16_t elas_s16x2 __attribute__((ext_vector_type(2)));

namespace {

__device__ __forceinline__ uint32_t lo_bytes(uint32_t x) { return x & 0x00ff00ffu; }
__device__ __forceinline__ uint32_t hi_bytes(uint32_t x) { return (x >> 8) & 0x00ff00ffu; }

// 128 + clamp(s >> 2, -128, 127) in both 16-bit halves
__device__ __forceinline__ uint32_t sobel_byte2(uint32_t s) {
    elas_s16x2 v = __builtin_bit_cast(elas_s16x2, s);
    v = v >> (elas_s16x2)(2);
    v = __builtin_elementwise_max(v, (elas_s16x2)(-128));
    v = __builtin_elementwise_min(v, (elas_s16x2)(127));
    v = v + (elas_s16x2)(128);
    return __builtin_bit_cast(uint32_t, v);
}

__device__ __forceinline__ uint32_t sub16x2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(elas_s16x2, a) - __builtin_bit_cast(elas_s16x2, b));
}

// bytes of (even, odd) halves -> [e.b0, o.b0, e.b2, o.b2]
__device__ __forceinline__ uint32_t pack_eo(uint32_t e, uint32_t o) { return __builtin_amdgcn_perm(o, e, 0x06020400u); }

constexpr uint32_t sel4(int b0, int b1, int b2, int b3) {
    return (uint32_t)b0 | ((uint32_t)b1 << 8) | ((uint32_t)b2 << 16) | ((uint32_t)b3 << 24);
}

}  // namespace

__global__ __launch_bounds__(256) void k_descriptor(ElasLaunch L) {
    __shared__ uint32_t s_img[DT_IR * DT_DW];
    __shared__ uint32_t s_du[DT_UR * DT_DW];
    __shared__ uint32_t s_dv[DT_VR * DT_DW];
    const int tid = threadIdx.x;
    constexpr int N = DT_IR * DT_DW, K = (N + 255) / 256;
    // tile t -> pair, image, origin
    struct Tile {
        const ELAS_CONST_AS ElasPair* P;
        int W, H, img, u0, v0;
    };
    auto tile_of = [&](int t) {
        Tile T;
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        T.P = &L.pairs[p];
        T.W = T.P->W;
        T.H = T.P->H;
        const int local = t - L.prefix[p];
        const int tx = (T.W + DT_W - 1) / DT_W, ty = (T.H + DT_H - 1) / DT_H;
        T.img = local / (tx * ty);
        const int r = local - T.img * tx * ty;
        T.u0 = (r % tx) * DT_W;
        T.v0 = (r / tx) * DT_H;
        return T;
    };
    // Image window of a tile: every byte load is issued unconditionally from
    // a clamped address (so they are all in flight together); out-of-image
    // bytes are zeroed when the bytes are packed into dwords for LDS.
    auto load_window = [&](const Tile& T, uint32_t (&b)[K][4]) {
        const uint8_t* I = T.P->img[T.img];
        #pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + 256 * k;
            const int rr = i / DT_DW, dw = i - rr * DT_DW;
            const int v = min(max(T.v0 - 3 + rr, 0), T.H - 1), ub = T.u0 - 8 + 4 * dw;
            #pragma unroll
            for (int j = 0; j < 4; j++) {
                const int u = min(max(ub + j, 0), T.W - 1);
                b[k][j] = elas_ld(I + (size_t)v * T.W + u);
            }
        }
    };
    auto store_window = [&](const Tile& T, const uint32_t (&b)[K][4]) {
        #pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + 256 * k;
            const int rr = i / DT_DW, dw = i - rr * DT_DW;
            const int v = T.v0 - 3 + rr, ub = T.u0 - 8 + 4 * dw;
            const bool rowok = i < N && v >= 0 && v < T.H;
            uint32_t x = 0;
            #pragma unroll
            for (int j = 0; j < 4; j++) {
                const int u = ub + j;
                x |= ((rowok && u >= 0 && u < T.W) ? b[k][j] : 0u) << (8 * j);
            }
            if (i < N) s_img[i] = x;
        }
    };
    // The next tile's window is loaded while this tile is computed and
    // stored, so the load latency (long under the descriptor write stream)
    // overlaps the work.
    uint32_t bn[K][4];
    int t = blockIdx.x;
    Tile Tn;
    if (t < L.total) {
        Tn = tile_of(t);
        load_window(Tn, bn);
    }
    for (; t < L.total; t += gridDim.x) {
        const Tile T = Tn;
        const auto& P = *T.P;
        const int W = T.W, H = T.H, img = T.img, u0 = T.u0, v0 = T.v0;
        store_window(T, bn);
        if (t + (int)gridDim.x < L.total) {
            Tn = tile_of(t + gridDim.x);
            load_window(Tn, bn);
        }
        elas_lds_barrier();
        // du (groups of four columns 4g..4g+3, g = 1..DT_NG) and dv
        for (int i = tid; i < (DT_UR + DT_VR) * DT_NG; i += 256) {
            const bool is_du = i < DT_UR * DT_NG;
            const int j = is_du ? i : i - DT_UR * DT_NG;
            const int row = j / DT_NG, g = j - row * DT_NG + 1;
            if (is_du) {
                // du row `row` (image row v) takes window rows row .. row + 2 (v-1 .. v+1)
                const uint32_t* a = s_img + row * DT_DW + g;
                const uint32_t* b = a + DT_DW;
                const uint32_t* c = b + DT_DW;
                const uint32_t Pa = a[-1], Oa = a[0], Na = a[1];
                const uint32_t Pb = b[-1], Ob = b[0], Nb = b[1];
                const uint32_t Pc = c[-1], Oc = c[0], Nc = c[1];
                const uint32_t Vlo = lo_bytes(Oa) + lo_bytes(Oc) + (lo_bytes(Ob) << 1);  // cols c, c+2
                const uint32_t Vhi = hi_bytes(Oa) + hi_bytes(Oc) + (hi_bytes(Ob) << 1);  // cols c+1, c+3
                const uint32_t Vm1 = (Pa >> 24) + (Pc >> 24) + ((Pb >> 24) << 1);         // col c-1
                const uint32_t Vp4 = (Na & 0xffu) + (Nc & 0xffu) + ((Nb & 0xffu) << 1);   // col c+4
                const uint32_t E = sub16x2(Vm1 | (Vhi << 16), Vhi);                       // S(c), S(c+2)
                const uint32_t O = sub16x2(Vlo, (Vlo >> 16) | (Vp4 << 16));               // S(c+1), S(c+3)
                s_du[row * DT_DW + g] = pack_eo(sobel_byte2(E), sobel_byte2(O));
            } else {
                // dv row `row` (image row v) takes window rows row + 1 (v-1) and row + 3 (v+1)
                uint32_t He[2], Ho[2];
                #pragma unroll
                for (int h = 0; h < 2; h++) {
                    const uint32_t* a = s_img + (row + 1 + 2 * h) * DT_DW + g;
                    const uint32_t Pw = a[-1], Ow = a[0], Nw = a[1];
                    const uint32_t lo = lo_bytes(Ow), hi = hi_bytes(Ow);
                    const uint32_t le = lo_bytes(__builtin_amdgcn_alignbyte(Ow, Pw, 3));  // cols c-1, c+1
                    const uint32_t ro = lo_bytes(__builtin_amdgcn_alignbyte(Nw, Ow, 2));  // cols c+2, c+4
                    He[h] = le + hi + (lo << 1);
                    Ho[h] = lo + ro + (hi << 1);
                }
                s_dv[row * DT_DW + g] = pack_eo(sobel_byte2(sub16x2(He[0], He[1])), sobel_byte2(sub16x2(Ho[0], Ho[1])));
            }
        }
        elas_lds_barrier();
        // four adjacent pixels per thread: row pr, columns 4 pg .. 4 pg + 3
        {
            const int pr = tid / DT_GP, pg = tid % DT_GP;
            const int G = pg + 2;
            const int v = v0 + pr;
            const uint32_t* du = s_du + pr * DT_DW + G;  // du rows v-2 .. v+2 = pr .. pr+4
            const uint32_t* dv = s_dv + pr * DT_DW + G;  // dv rows v-1 .. v+1 = pr .. pr+2
            const uint32_t um2 = du[0];
            const uint32_t um1p = du[DT_DW - 1], um1 = du[DT_DW], um1n = du[DT_DW + 1];
            const uint32_t u0p = du[2 * DT_DW - 1], u00 = du[2 * DT_DW], u0n = du[2 * DT_DW + 1];
            const uint32_t up1p = du[3 * DT_DW - 1], up1 = du[3 * DT_DW], up1n = du[3 * DT_DW + 1];
            const uint32_t up2 = du[4 * DT_DW];
            const uint32_t vm1 = dv[0];
            const uint32_t v0p = dv[DT_DW - 1], v00 = dv[DT_DW], v0n = dv[DT_DW + 1];
            const uint32_t vp1 = dv[2 * DT_DW];
            uint8_t* D = P.desc[img];
            #pragma unroll
            for (int k = 0; k < 4; k++) {
                // a pixel's columns k-2 .. k+2 lie in (prev, own) for k < 2, in (own, next) otherwise
                const int b = k < 2 ? 4 + k : k;
                const uint32_t wl_m1 = k < 2 ? um1p : um1, wh_m1 = k < 2 ? um1 : um1n;
                const uint32_t wl_0 = k < 2 ? u0p : u00, wh_0 = k < 2 ? u00 : u0n;
                const uint32_t wl_p1 = k < 2 ? up1p : up1, wh_p1 = k < 2 ? up1 : up1n;
                const uint32_t wl_v = k < 2 ? v0p : v00, wh_v = k < 2 ? v00 : v0n;
                uint4 q;
                q.x = __builtin_amdgcn_perm(um2, __builtin_amdgcn_perm(wh_m1, wl_m1, sel4(12, b - 2, b, b + 2)),
                                            sel4(4 + k, 1, 2, 3));
                q.y = __builtin_amdgcn_perm(wh_0, wl_0, sel4(b - 1, b, b, b + 1));
                q.z = __builtin_amdgcn_perm(up2, __builtin_amdgcn_perm(wh_p1, wl_p1, sel4(b - 2, b, b + 2, 12)),
                                            sel4(0, 1, 2, 4 + k));
                const uint32_t w = __builtin_amdgcn_perm(wh_v, wl_v, sel4(12, b - 1, b + 1, 12));
                q.w = __builtin_amdgcn_perm(vp1, __builtin_amdgcn_perm(vm1, w, sel4(4 + k, 1, 2, 3)), sel4(0, 1, 2, 4 + k));
                const int u = u0 + 4 * pg + k;
                if (!(u >= 3 && u <= W - 4 && v >= 3 && v <= H - 4)) q = make_uint4(0u, 0u, 0u, 0u);
                if (u < W && v < H) elas_st(reinterpret_cast<uint4*>(D + ((size_t)v * W + u) * 16), q);
            }
        }
        elas_lds_barrier();
    }
}

void elas_launch_descriptor(const ElasLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_descriptor, dim3(grid), dim3(256), 0, s, L);
}
