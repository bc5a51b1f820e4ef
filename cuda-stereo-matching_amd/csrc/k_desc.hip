// k_desc.hip -- fused Sobel + 16-byte descriptor (gfx950).
//
// Restates src/Kernels/KerDescriptor.cu:9-100 (kerComputeSobel and
// kerComputeDescriptor) as one kernel.  du/dv never leave registers:
// algorithmic traffic is 1 B read + 16 B written per pixel.
//
// du = 128 + clamp(S_u >> 2, -128, 127) is the integer form of the
// reference's (uint8_t)(min(max(-128.f, S/4), 127.f) + 128)
// (KerDescriptor.cu:40-41): S/4 is exact in float and truncating the
// non-negative sum is floor.  S_u(u) = V(u-1) - V(u+1) with the vertical
// weighted sum V = I(v-1) + 2 I(v) + I(v+1); S_v(v) = H(v-1) - H(v+1) with the
// horizontal one H = I(u-1) + 2 I(u) + I(u+1).  Descriptor bytes outside
// u in [3, W-4], v in [3, H-4] are written as 0 (the reference leaves them
// unwritten).  Byte order (KerDescriptor.cu:76-92), for row y, column u:
//   du(y-2,u) du(y-1,u-2) du(y-1,u) du(y-1,u+2) | du(y,u-1) du(y,u) du(y,u) du(y,u+1) |
//   du(y+1,u-2) du(y+1,u) du(y+1,u+2) du(y+2,u) | dv(y-1,u) dv(y,u-1) dv(y,u+1) dv(y+1,u)
//
// One pixel column per lane.  A wave owns a strip of 64 consecutive columns
// u = x0 - 3 + lane and walks down the image rows y0-3 .. y0+SH+2 of a band
// (SH = 6, 12 or 18 output rows, per pair: short strips give a small image
// more waves);
// lanes 3 .. 60 are the strip's ELAS_DESC_SW output columns, the three lanes
// on either side only feed their neighbours.  Per image row the wave computes
// H of that row and V, du, dv of the row above it, for its own columns;
// values of neighbouring columns come from cross-lane permutes.  The last
// five du rows and three dv rows stay in registers, so every pixel's
// descriptor is assembled with a handful of byte permutes, and the wave's
// 16-byte stores of one row cover consecutive pixels (928 contiguous bytes
// per store instruction).  No LDS, no barriers: the four waves of a
// workgroup work on independent strips.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "elas_common.h"
#include "elas_device.h"

static_assert(ELAS_DESC_SW == 58, "lanes 3 .. 60 are the output columns");

typedef int16_t elas_s16x2 __attribute__((ext_vector_type(2)));

namespace {

// value of the lane to the left (u - 1) / right (u + 1) through the LDS
// crossbar (ds_bpermute, no LDS memory); lanes 0 / 63 read lanes 63 / 0,
// values no output lane uses.  ixl / ixr: byte addresses 4 (lane -/+ 1).
struct Xlane {
    int ixl, ixr;
    __device__ __forceinline__ int left(int x) const { return __builtin_amdgcn_ds_bpermute(ixl, x); }
    __device__ __forceinline__ int right(int x) const { return __builtin_amdgcn_ds_bpermute(ixr, x); }
    __device__ __forceinline__ uint32_t left(uint32_t x) const { return (uint32_t)left((int)x); }
    __device__ __forceinline__ uint32_t right(uint32_t x) const { return (uint32_t)right((int)x); }
};

// 128 + clamp(s >> 2, -128, 127) in both 16-bit halves
__device__ __forceinline__ uint32_t sobel_byte2(uint32_t s) {
    elas_s16x2 v = __builtin_bit_cast(elas_s16x2, s);
    v = v >> (elas_s16x2)(2);
    v = __builtin_elementwise_max(v, (elas_s16x2)(-128));
    v = __builtin_elementwise_min(v, (elas_s16x2)(127));
    v = v + (elas_s16x2)(128);
    return __builtin_bit_cast(uint32_t, v);
}

__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) {
    return __builtin_amdgcn_perm(hi, lo, sel);
}

}  // namespace

// One strip: columns u = x0 - 3 + lane, output rows [y0, y1).  EDGE: the
// strip holds pixels whose descriptor is 0 (image sides, top or bottom rows);
// interior strips skip that test.
template <int SH, bool EDGE>
__device__ __forceinline__ void desc_strip(const ELAS_CONST_AS ElasPair& P, int img, int x0, int y0, int y1) {
    const int lane = threadIdx.x & 63;
    const Xlane X{((lane + 63) & 63) * 4, ((lane + 1) & 63) * 4};
    const int W = P.W, H = P.H;
    const int u = x0 - 3 + lane;
    // image bytes through a buffer resource: the row offset is scalar, the
    // column a per-lane constant
    const __amdgpu_buffer_rsrc_t I = elas_rsrc((void*)P.img[img], (uint32_t)W * (uint32_t)H);
    const int uc = min(max(u, 0), W - 1);
    auto load = [&](int R) -> uint32_t {
        return __builtin_amdgcn_raw_buffer_load_b8(I, uc, min(max(R, 0), H - 1) * W, 0);
    };
    // stores: byte offset 16 (y W + u); lanes and rows with nothing to store
    // take an offset past the buffer (saturating add), which the hardware
    // drops, so every store issues without a branch
    const uint32_t range = (uint32_t)W * (uint32_t)H * 16u;  // W * H <= 2^27
    const __amdgpu_buffer_rsrc_t D = elas_rsrc(P.desc[img], range);
    const uint32_t col = (lane >= 3 && lane < 3 + ELAS_DESC_SW && u < W) ? 16u * (uint32_t)u : 0xfffffff0u;
    const bool zcol = u < 3 || u > W - 4;

    // every image byte of the band is loaded before the first store: the
    // vector memory counter completes in issue order, so a load issued after
    // stores could only be waited for together with their write-back
    constexpr int DS_STEPS = SH + 6;  // image rows walked per band
    uint32_t q[DS_STEPS];
    #pragma unroll
    for (int k = 0; k < DS_STEPS; k++) q[k] = load(y0 - 3 + k);
    // state: image bytes and H of the two previous rows; packed du/dv (du in
    // byte 0, dv in byte 2) of the last four du rows, and the neighbourhood
    // words A (du at u-2, u, u+2), B (du at u-1, u, u, u+1) and D (dv at u-1,
    // u+1) of the last rows
    int i1 = 0, i2 = 0, h1 = 0, h2 = 0;
    uint32_t c1 = 0, c2 = 0, c3 = 0, c4 = 0, A1 = 0, A2 = 0, A3 = 0, B1 = 0, B2 = 0, D1 = 0, D2 = 0;
    #pragma unroll
    for (int j = 0; j < DS_STEPS; j++) {
        // image row R = y0 - 3 + j gives du/dv of row R - 1 and the
        // descriptor of row y = R - 3 (du rows y-2 .. y+2 = c4 .. c,
        // dv rows y-1 .. y+1 = c3 .. c1)
        const int iR = (int)q[j];
        const int hR = X.left(iR) + X.right(iR) + 2 * iR;
        const int V = i2 + 2 * i1 + iR;
        const int du = X.left(V) - X.right(V);
        const int dv = h2 - hR;
        const uint32_t c = sobel_byte2(perm((uint32_t)dv, (uint32_t)du, 0x05040100u));
        const uint32_t cl = X.left(c), cr = X.right(c);
        const uint32_t cl2 = X.left(cl), cr2 = X.right(cr);
        const uint32_t A = (cr2 << 16) | perm(c, cl2, 0x0c0c0400u);  // du u-2, u, u+2
        const uint32_t B = (cr << 24) | perm(c, cl, 0x0c040400u);    // du u-1, u, u, u+1
        const uint32_t Dn = perm(cr, cl, 0x0c0c0602u);               // dv u-1, u+1
        const int y = y0 - 6 + j;
        uint4 o;
        o.x = perm(A3, c4, 0x06050400u);
        o.y = B2;
        o.z = (c << 24) | A1;
        o.w = perm(c1, perm(D2, c3, 0x0c050402u), 0x06020100u);
        if (EDGE && (zcol || y < 3 || y > H - 4)) o = make_uint4(0u, 0u, 0u, 0u);
        if (j >= 6) {  // the first six rows only fill the registers
            const uint32_t row = y < y1 ? (uint32_t)y * (uint32_t)W * 16u : 0xfffffff0u;
            elas_st16_ranged(D, __builtin_elementwise_add_sat(row, col), o);
        }
        i2 = i1; i1 = iR; h2 = h1; h1 = hR;
        c4 = c3; c3 = c2; c2 = c1; c1 = c;
        A3 = A2; A2 = A1; A1 = A;
        B2 = B1; B1 = B;
        D2 = D1; D1 = Dn;
    }
}

template <int SH>
__global__ __launch_bounds__(256) void k_descriptor(ElasLaunch L) {
    const int stride = gridDim.x * 4;
    for (int t0 = blockIdx.x * 4 + (threadIdx.x >> 6); t0 < L.total; t0 += stride) {
        const int t = __builtin_amdgcn_readfirstlane(t0);  // one strip per wave
        const int p = elas_find_pair(L.prefix, L.npairs, t);
        const auto& P = L.pairs[p];
        const int W = P.W, H = P.H;
        const int ns = (W + ELAS_DESC_SW - 1) / ELAS_DESC_SW, nb = (H + SH - 1) / SH;
        const int local = t - L.prefix[p];
        const int img = local / (ns * nb);
        const int r = local - img * ns * nb;
        const int x0 = (r % ns) * ELAS_DESC_SW, y0 = (r / ns) * SH;
        const int y1 = min(y0 + SH, H);
        // output columns [x0, x0 + 58) meet u < 3 or u > W - 4, or rows
        // [y0, y1) meet y < 3 or y > H - 4
        if (x0 < 3 || x0 + ELAS_DESC_SW > W - 3 || y0 < 3 || y1 > H - 3)
            desc_strip<SH, true>(P, img, x0, y0, y1);
        else
            desc_strip<SH, false>(P, img, x0, y0, y1);
    }
}

void elas_launch_descriptor(const ElasLaunch& L, int sh, int grid, hipStream_t s) {
    grid = std::min(grid, (L.total + 3) / 4);  // four strips per workgroup
    if (sh == 6) hipLaunchKernelGGL(k_descriptor<6>, dim3(grid), dim3(256), 0, s, L);
    else if (sh == 12) hipLaunchKernelGGL(k_descriptor<12>, dim3(grid), dim3(256), 0, s, L);
    else hipLaunchKernelGGL(k_descriptor<18>, dim3(grid), dim3(256), 0, s, L);
}
