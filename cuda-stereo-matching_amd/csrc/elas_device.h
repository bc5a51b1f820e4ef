// elas_device.h -- device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "elas_common.h"

// Pair owning flattened tile t: the last p with prefix[p] <= t.  t is
// workgroup-uniform, so the search runs on the scalar unit.
__device__ __forceinline__ int elas_find_pair(const ELAS_CONST_AS int32_t* prefix, int np, int t) {
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Loads through pointers known to address global memory (the ElasPair
// pointers are generic; without the cast the compiler emits flat loads that
// also wait on the LDS counter).
#define ELAS_AS_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T elas_ld(const T* p) {
    return *(const ELAS_AS_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ void elas_st(T* p, const T& v) {
    *(ELAS_AS_GLOBAL T*)p = v;
}
typedef uint32_t elas_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 elas_ld(const uint4* p) {
    const elas_u32x4 v = *(const ELAS_AS_GLOBAL elas_u32x4*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void elas_st(uint4* p, const uint4& v) {
    elas_u32x4 x;
    x.x = v.x; x.y = v.y; x.z = v.z; x.w = v.w;
    *(ELAS_AS_GLOBAL elas_u32x4*)p = x;
}

typedef uint32_t elas_u32x4 __attribute__((ext_vector_type(4)));

// Buffer resource over [base, base + bytes) for raw buffer stores (stride 0,
// the gfx9 default format word).  bytes < 2^32.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t elas_rsrc(void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)bytes, 0x00020000);
}

// 16-byte store at byte offset off of a buffer resource.  An access at or
// beyond the resource's range is dropped by the hardware, so a lane with
// nothing to store passes the range size as its offset instead of branching
// around the store: the store then always issues, and the compiler's count
// of outstanding memory operations stays exact across it.
__device__ __forceinline__ void elas_st16_ranged(__amdgpu_buffer_rsrc_t r, uint32_t off, uint4 q) {
    elas_u32x4 v = {q.x, q.y, q.z, q.w};
    __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)off, 0, 0);
}

// Offset past any buffer of this library (every range is below 2^31 bytes):
// a load there returns 0 and a store there is dropped.
#define ELAS_OOR 0xfffffff0u

// 16- and 4-byte buffer loads at byte offset off (ELAS_OOR: 0, no access),
// and a 4-byte buffer store (ELAS_OOR: dropped).  Issued without a branch,
// they keep the compiler's count of outstanding memory operations exact, so
// a wait for one load need not wait for everything issued before it.
__device__ __forceinline__ uint4 elas_ld16_ranged(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const elas_u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t elas_ld4_ranged(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
__device__ __forceinline__ void elas_st4_ranged(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t v) {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)off, 0, 0);
}

// Workgroup barrier that orders LDS only.  __syncthreads() is also a
// workgroup memory fence, which waits for every global load and store the
// wave has in flight (vmcnt 0): a wave would stall at each tile boundary until
// its result stores are written back.  The kernels of this library exchange
// data between the threads of a workgroup only through LDS (global data a
// workgroup reads is its own or was written by an earlier launch), so here a
// wave's prefetched loads and finished tile's stores stay in flight across
// the barrier.  The asm's memory clobber keeps the compiler from moving
// memory accesses across it.
__device__ __forceinline__ void elas_lds_barrier() {
    __asm__ __volatile__("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Sum of |a_i - b_i| over 16 bytes, accumulated onto acc (4 x v_sad_u8).
__device__ __forceinline__ uint32_t elas_sad16(const uint4 a, const uint4 b, uint32_t acc) {
    acc = __builtin_amdgcn_sad_u8(a.x, b.x, acc);
    acc = __builtin_amdgcn_sad_u8(a.y, b.y, acc);
    acc = __builtin_amdgcn_sad_u8(a.z, b.z, acc);
    acc = __builtin_amdgcn_sad_u8(a.w, b.w, acc);
    return acc;
}

// c + (sum of |a_i - b_i| over 16 bytes) << 16 (4 x v_sad_hi_u8): a key whose
// low 16 bits, taken from c, identify the candidate
__device__ __forceinline__ uint32_t elas_sad16_hi(const uint4 a, const uint4 b, uint32_t c) {
    c = __builtin_amdgcn_sad_hi_u8(a.x, b.x, c);
    c = __builtin_amdgcn_sad_hi_u8(a.y, b.y, c);
    c = __builtin_amdgcn_sad_hi_u8(a.z, b.z, c);
    c = __builtin_amdgcn_sad_hi_u8(a.w, b.w, c);
    return c;
}

// float -> int32 / uint32 with saturation and NaN -> 0, written out so the
// result does not depend on how the compiler treats out-of-range casts: the
// cast only sees values clamped into range (2147483520 and 4294967040 are
// the largest floats below 2^31 and 2^32), and selects, not branches, pick
// the saturated ends.
__device__ __forceinline__ int32_t elas_f2i_sat(float x) {
    int32_t r = (int32_t)fminf(fmaxf(x, -2147483648.0f), 2147483520.0f);
    r = x >= 2147483648.0f ? INT32_MAX : r;
    return x != x ? 0 : r;
}
__device__ __forceinline__ uint32_t elas_f2u_sat(float x) {
    uint32_t r = (uint32_t)fminf(fmaxf(x, 0.0f), 4294967040.0f);
    r = x >= 4294967296.0f ? 0xffffffffu : r;
    return x != x ? 0u : r;
}

// floor(x / g) for g > 0 (the reference evaluates floor((float)x / g), equal
// for |x| < 2^20).
__device__ __forceinline__ int elas_floordiv(int x, int g) {
    return x >= 0 ? x / g : -((-x + g - 1) / g);
}
