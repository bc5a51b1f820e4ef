// elas_triangle.cpp -- host Delaunay bridge over the reference's Triangle.
//
// The reference triangulates on the host with TRILIB::triangulate("zQB")
// (src/Elas.cu:282-328).  Triangle keeps its state in namespace-scope
// globals (src/triangle.cpp:545-554, the RNG reset at :4034), so one instance
// is not reentrant.  The build compiles the reference's triangle.cpp
// unmodified ELAS_TRI_INSTANCES times with -DTRILIB=elas_triK, each inside
// elas_tri_unit.cpp, which adds a driver for a precomputed vertex order: each
// copy gets its own namespace and globals, and worker thread K of the pool below uses
// only copy K.  Meshes are identical to a single serial Triangle because every
// call starts from the same reset state.
#include <algorithm>
#include <atomic>
#include <immintrin.h>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <sched.h>
#include <mutex>
#include <thread>
#include <vector>

#include "elas_triangle.h"

#define TRILIB elas_tri0
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri1
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri2
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri3
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri4
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri5
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri6
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri7
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri8
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri9
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri10
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri11
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri12
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri13
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri14
#include "triangle.h"
#undef TRILIB
#define TRILIB elas_tri15
#include "triangle.h"
#undef TRILIB

#define ELAS_TRI_DECL(NS)                                                                             \
    namespace NS {                                                                                    \
    void* elas_tri_setup(const float* pts, int n);                                                    \
    int elas_tri_finish(void* setup, ElasTriOrder* order, bool integer, int depth, ElasParallelFor par, \
                        void* parctx, std::vector<int32_t>* out, int32_t* dst, int64_t cap, bool dst16); \
    }
ELAS_TRI_DECL(elas_tri0) ELAS_TRI_DECL(elas_tri1) ELAS_TRI_DECL(elas_tri2) ELAS_TRI_DECL(elas_tri3)
ELAS_TRI_DECL(elas_tri4) ELAS_TRI_DECL(elas_tri5) ELAS_TRI_DECL(elas_tri6) ELAS_TRI_DECL(elas_tri7)
ELAS_TRI_DECL(elas_tri8) ELAS_TRI_DECL(elas_tri9) ELAS_TRI_DECL(elas_tri10) ELAS_TRI_DECL(elas_tri11)
ELAS_TRI_DECL(elas_tri12) ELAS_TRI_DECL(elas_tri13) ELAS_TRI_DECL(elas_tri14) ELAS_TRI_DECL(elas_tri15)

namespace {

typedef void* (*SetupFn)(const float*, int);
const SetupFn kSetup[ELAS_TRI_INSTANCES] = {
    elas_tri0::elas_tri_setup,  elas_tri1::elas_tri_setup,  elas_tri2::elas_tri_setup,  elas_tri3::elas_tri_setup,
    elas_tri4::elas_tri_setup,  elas_tri5::elas_tri_setup,  elas_tri6::elas_tri_setup,  elas_tri7::elas_tri_setup,
    elas_tri8::elas_tri_setup,  elas_tri9::elas_tri_setup,  elas_tri10::elas_tri_setup, elas_tri11::elas_tri_setup,
    elas_tri12::elas_tri_setup, elas_tri13::elas_tri_setup, elas_tri14::elas_tri_setup, elas_tri15::elas_tri_setup};

typedef int (*FinishFn)(void*, ElasTriOrder*, bool, int, ElasParallelFor, void*, std::vector<int32_t>*, int32_t*,
                        int64_t, bool);
const FinishFn kFinish[ELAS_TRI_INSTANCES] = {
    elas_tri0::elas_tri_finish,  elas_tri1::elas_tri_finish,  elas_tri2::elas_tri_finish,
    elas_tri3::elas_tri_finish,  elas_tri4::elas_tri_finish,  elas_tri5::elas_tri_finish,
    elas_tri6::elas_tri_finish,  elas_tri7::elas_tri_finish,  elas_tri8::elas_tri_finish,
    elas_tri9::elas_tri_finish,  elas_tri10::elas_tri_finish, elas_tri11::elas_tri_finish,
    elas_tri12::elas_tri_finish, elas_tri13::elas_tri_finish, elas_tri14::elas_tri_finish,
    elas_tri15::elas_tri_finish};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ---- block cache ---------------------------------------------------------------

namespace {

#define ELAS_BLOCK_MIN_SHIFT 16                   // 64 KiB
#define ELAS_BLOCK_CLASSES 16                     // up to 2 GiB
#define ELAS_BLOCK_CACHE_MAX ((size_t)512 << 20)  // bytes kept for reuse, shared
#define ELAS_BLOCK_LOCAL_MAX ((size_t)32 << 20)   // bytes kept for reuse, per thread
#define ELAS_BLOCK_HEADER 64                      // size class in front of a block (keeps 16-B alignment)

struct BlockCache {
    std::mutex m;
    std::vector<void*> free_[ELAS_BLOCK_CLASSES];
    size_t cached = 0;
};

BlockCache& block_cache() {
    static BlockCache* c = new BlockCache();  // lives until process exit
    return *c;
}

size_t class_bytes(int cls) { return (size_t)1 << (ELAS_BLOCK_MIN_SHIFT + cls); }

// Blocks a thread freed go to its own list first: the thread takes them back
// without a lock, and mostly while their lines are still in its cache.  At
// thread exit they move to the shared cache (or back to malloc).
struct LocalCache {
    std::vector<void*> free_[ELAS_BLOCK_CLASSES];
    size_t cached = 0;
    ~LocalCache() {
        BlockCache& c = block_cache();
        std::lock_guard<std::mutex> g(c.m);
        for (int cls = 0; cls < ELAS_BLOCK_CLASSES; cls++)
            for (void* p : free_[cls]) {
                if (c.cached + class_bytes(cls) <= ELAS_BLOCK_CACHE_MAX) {
                    c.free_[cls].push_back(p);
                    c.cached += class_bytes(cls);
                } else {
                    free(p);
                }
            }
    }
};

LocalCache& local_cache() {
    thread_local LocalCache c;
    return c;
}

}  // namespace

void* elas_tri_block_alloc(size_t bytes) {
    if (bytes + ELAS_BLOCK_HEADER < ((size_t)1 << ELAS_BLOCK_MIN_SHIFT)) {
        char* p = (char*)malloc(bytes + ELAS_BLOCK_HEADER);
        if (!p) return nullptr;
        *(int*)p = -1;
        return p + ELAS_BLOCK_HEADER;
    }
    int cls = 0;
    while (class_bytes(cls) < bytes + ELAS_BLOCK_HEADER) cls++;
    if (cls >= ELAS_BLOCK_CLASSES) return nullptr;  // Triangle reports out of memory
    char* p = nullptr;
    LocalCache& lc = local_cache();
    if (!lc.free_[cls].empty()) {
        p = (char*)lc.free_[cls].back();
        lc.free_[cls].pop_back();
        lc.cached -= class_bytes(cls);
    } else {
        BlockCache& c = block_cache();
        std::lock_guard<std::mutex> g(c.m);
        if (!c.free_[cls].empty()) {
            p = (char*)c.free_[cls].back();
            c.free_[cls].pop_back();
            c.cached -= class_bytes(cls);
        }
    }
    if (!p) p = (char*)malloc(class_bytes(cls));
    if (!p) return nullptr;
    *(int*)p = cls;
    return p + ELAS_BLOCK_HEADER;
}

void elas_tri_block_free(void* ptr) {
    if (!ptr) return;
    char* p = (char*)ptr - ELAS_BLOCK_HEADER;
    const int cls = *(int*)p;
    if (cls >= 0) {
        const size_t sz = class_bytes(cls);
        LocalCache& lc = local_cache();
        if (lc.cached + sz <= ELAS_BLOCK_LOCAL_MAX) {
            lc.free_[cls].push_back(p);
            lc.cached += sz;
            return;
        }
        BlockCache& c = block_cache();
        std::lock_guard<std::mutex> g(c.m);
        if (c.cached + sz <= ELAS_BLOCK_CACHE_MAX) {
            c.free_[cls].push_back(p);
            c.cached += sz;
            return;
        }
    }
    free(p);
}

void ElasTriOrder::resize(int n, bool narrow_entries) {
    narrow = narrow_entries;
    lx.resize(n);
    ly.resize(n);
    arr.resize(n);
    rank.resize(n);
    if (narrow) {
        qx.resize(n);
        qy.resize(n);
        ql.resize(n);
        qr.resize(n);
    } else {
        px.resize(n);
        py.resize(n);
        tl.resize(n);
        tr.resize(n);
    }
}

void ElasTriOrder::pack() {
    const int n = (int)lx.size();
    for (int k = 0; k < n; k++) rank[ly[k]] = k;
    for (int k = 0; k < n; k++) px[k] = (uint64_t)(uint32_t)k << 32 | (uint32_t)rank[lx[k]];
    for (int k = 0; k < n; k++) rank[lx[k]] = k;
    for (int k = 0; k < n; k++) py[k] = (uint64_t)(uint32_t)rank[ly[k]] << 32 | (uint32_t)k;
}

namespace {

// cut() over entries of type E: (x rank << H | y rank), H = half E's bits
template <typename E>
void cut_entries(E* px, E* py, E* tl, E* tr, const int32_t* ly, int32_t* arr, int lo, int size, int axis,
                 int levels) {
    constexpr int H = (int)sizeof(E) * 4;
    constexpr E LOW = (E)(((E)1 << H) - 1);
    if (size <= 3) {
        for (int k = 0; k < size; k++) arr[lo + k] = ly[px[lo + k] & LOW];
        return;
    }
    if (levels == 0) return;
    const int divider = size >> 1;
    const E* cutl = axis ? py : px;
    E* other = axis ? px : py;
    const int sh = axis ? 0 : H;
    const E last = (E)((cutl[lo + divider - 1] >> sh) & LOW);  // rank of the left's last point
    // stable split of the other list without branches (the side of a point
    // is not predictable): every entry is stored to both sides and only its
    // own side's cursor moves on.  The store past a side's end lands inside
    // the node's range of tl / tr: a < divider and b < size - divider until
    // the last entry.
    E* l = tl + lo;
    E* r = tr + lo;
    const E* o = other + lo;
    int a = 0, b = 0;
    for (int k = 0; k < size; k++) {
        const E e = o[k];
        const int right = (E)((e >> sh) & LOW) > last;
        l[a] = e;
        r[b] = e;
        a += 1 - right;
        b += right;
    }
    memcpy(other + lo, l, sizeof(E) * (size_t)divider);
    memcpy(other + lo + divider, r, sizeof(E) * (size_t)(size - divider));
    cut_entries(px, py, tl, tr, ly, arr, lo, divider, 1 - axis, levels - 1);
    cut_entries(px, py, tl, tr, ly, arr, lo + divider, size - divider, 1 - axis, levels - 1);
}

}  // namespace

void ElasTriOrder::cut(int lo, int size, int axis, int levels) {
    if (narrow)
        cut_entries<uint32_t>(qx.data(), qy.data(), ql.data(), qr.data(), ly.data(), arr.data(), lo, size, axis,
                              levels);
    else
        cut_entries<uint64_t>(px.data(), py.data(), tl.data(), tr.data(), ly.data(), arr.data(), lo, size, axis,
                              levels);
}

// The integer keys of n points [x0, y0, ...]: kx, ky, their bounds in
// lohi (min x, max x, min y, max y); returns nonzero when a coordinate is
// not an integer of magnitude below 2^24 (NaN included).  Eight points per
// step with AVX2 (the library's Triangle copies are built for x86-64-v3), the
// rest one by one; a coordinate out of range converts as 0 and fails the
// exactness test.
static __attribute__((target("avx2"))) int convert_keys(const float* p, int n, int32_t* kx, int32_t* ky, int* lohi) {
    int i = 0, bad = 0;
    int lox = INT32_MAX, hix = INT32_MIN, loy = INT32_MAX, hiy = INT32_MIN;
    if (n >= 8) {
        const __m256 lim = _mm256_set1_ps(16777216.0f);
        const __m256 sign = _mm256_set1_ps(-0.0f);
        __m256i vlox = _mm256_set1_epi32(INT32_MAX), vhix = _mm256_set1_epi32(INT32_MIN);
        __m256i vloy = vlox, vhiy = vhix;
        __m256 vbad = _mm256_setzero_ps();
        for (; i + 8 <= n; i += 8) {
            const __m256 a = _mm256_loadu_ps(p + 2 * i), b = _mm256_loadu_ps(p + 2 * i + 8);
            // (x0 y0 x1 y1 | x2 y2 x3 y3), (x4 y4 x5 y5 | x6 y6 x7 y7) -> x0..x7, y0..y7
            const __m256 xs = _mm256_castpd_ps(_mm256_permute4x64_pd(_mm256_castps_pd(_mm256_shuffle_ps(a, b, 0x88)), 0xD8));
            const __m256 ys = _mm256_castpd_ps(_mm256_permute4x64_pd(_mm256_castps_pd(_mm256_shuffle_ps(a, b, 0xDD)), 0xD8));
            const __m256 okx = _mm256_cmp_ps(_mm256_andnot_ps(sign, xs), lim, _CMP_LT_OQ);
            const __m256 oky = _mm256_cmp_ps(_mm256_andnot_ps(sign, ys), lim, _CMP_LT_OQ);
            const __m256i ix = _mm256_cvttps_epi32(_mm256_and_ps(xs, okx));
            const __m256i iy = _mm256_cvttps_epi32(_mm256_and_ps(ys, oky));
            vbad = _mm256_or_ps(vbad, _mm256_cmp_ps(_mm256_cvtepi32_ps(ix), xs, _CMP_NEQ_UQ));
            vbad = _mm256_or_ps(vbad, _mm256_cmp_ps(_mm256_cvtepi32_ps(iy), ys, _CMP_NEQ_UQ));
            _mm256_storeu_si256((__m256i*)(kx + i), ix);
            _mm256_storeu_si256((__m256i*)(ky + i), iy);
            vlox = _mm256_min_epi32(vlox, ix);
            vhix = _mm256_max_epi32(vhix, ix);
            vloy = _mm256_min_epi32(vloy, iy);
            vhiy = _mm256_max_epi32(vhiy, iy);
        }
        bad = _mm256_movemask_ps(vbad);
        alignas(32) int32_t t[4][8];
        _mm256_store_si256((__m256i*)t[0], vlox);
        _mm256_store_si256((__m256i*)t[1], vhix);
        _mm256_store_si256((__m256i*)t[2], vloy);
        _mm256_store_si256((__m256i*)t[3], vhiy);
        for (int k = 0; k < 8; k++) {
            lox = std::min(lox, t[0][k]);
            hix = std::max(hix, t[1][k]);
            loy = std::min(loy, t[2][k]);
            hiy = std::max(hiy, t[3][k]);
        }
    }
    for (; i < n; i++) {
        const float x = p[2 * i], y = p[2 * i + 1];
        const int ix = std::fabs(x) < 16777216.0f ? (int)x : 0, iy = std::fabs(y) < 16777216.0f ? (int)y : 0;
        bad |= ((float)ix != x) | ((float)iy != y);
        kx[i] = ix;
        ky[i] = iy;
        lox = std::min(lox, ix);
        hix = std::max(hix, ix);
        loy = std::min(loy, iy);
        hiy = std::max(hiy, iy);
    }
    lohi[0] = lox;
    lohi[1] = hix;
    lohi[2] = loy;
    lohi[3] = hiy;
    return bad;
}

// Whether keys come in y order (bit 0) and in strictly increasing (y, x)
// order (bit 1); eight neighbour pairs per step with AVX2.
static __attribute__((target("avx2"))) int input_order(const int32_t* kx, const int32_t* ky, int n) {
    int i = 1;
    __m256i ydown = _mm256_setzero_si256(), notyx = _mm256_setzero_si256();
    for (; i + 8 <= n; i += 8) {
        const __m256i y1 = _mm256_loadu_si256((const __m256i*)(ky + i)), y0 = _mm256_loadu_si256((const __m256i*)(ky + i - 1));
        const __m256i x1 = _mm256_loadu_si256((const __m256i*)(kx + i)), x0 = _mm256_loadu_si256((const __m256i*)(kx + i - 1));
        const __m256i down = _mm256_cmpgt_epi32(y0, y1);  // y falls
        const __m256i same = _mm256_cmpeq_epi32(y0, y1);
        const __m256i xnotup = _mm256_xor_si256(_mm256_cmpgt_epi32(x1, x0), _mm256_set1_epi32(-1));
        ydown = _mm256_or_si256(ydown, down);
        notyx = _mm256_or_si256(notyx, _mm256_or_si256(down, _mm256_and_si256(same, xnotup)));
    }
    bool ys = _mm256_testz_si256(ydown, ydown), yxs = _mm256_testz_si256(notyx, notyx);
    for (; i < n; i++) {
        const int dy = ky[i] - ky[i - 1];
        ys = ys && dy >= 0;
        yxs = yxs && ((dy > 0) || (dy == 0 && kx[i] > kx[i - 1]));
    }
    return (ys ? 1 : 0) | (yxs ? 2 : 0);
}

// Support points have integer coordinates, so both sort orders come from
// counting sorts, in O(n) instead of Triangle's randomised quicksort and
// quickselect, which take about half of a support-point mesh's serial time.
// The (y, x) order comes first: support points come row by row (y
// non-decreasing), the left image's also in (y, x) order, the right image's
// nearly so (a row's points move by their disparities), so it is the input
// order, or the input with each row sorted by insertion, or (any other
// input, or rows far from sorted) two counting sorts.  One stable counting
// sort of that order by x then gives the (x, y) order, each point's x rank
// and, with its y rank k, both packed entries.  Equal points are neighbours
// in (y, x) order.
namespace {

// a thread's scratch for the vertex order of the meshes it runs
struct OrderScratch {
    std::vector<int32_t> cnt, kxv, kyv, tmp;
    ElasTriOrder order;
};
OrderScratch& order_scratch() {
    thread_local OrderScratch s;
    return s;
}

}  // namespace

bool elas_tri_order_integer(const float* p, int n, ElasTriOrder& o) {
    if (n < 3) return false;
    OrderScratch& sc = order_scratch();
    std::vector<int32_t>&cnt = sc.cnt, &kxv = sc.kxv, &kyv = sc.kyv, &tmp = sc.tmp;
    kxv.resize(n);
    kyv.resize(n);
    int32_t* kx = kxv.data();
    int32_t* ky = kyv.data();
    // integer coordinates below 2^24 in magnitude
    int lohi[4];
    int bad = convert_keys(p, n, kx, ky, lohi);
    const int lox = lohi[0], hix = lohi[1], loy = lohi[2], hiy = lohi[3];
    if (bad) return false;
    for (int k = 0; k < 4; k++) o.lohi[k] = lohi[k];
    int sorted = input_order(kx, ky, n);
    const bool ysorted = sorted & 1, yxsorted = sorted & 2;
    const long rx = (long)hix - lox + 1, ry = (long)hiy - loy + 1;
    if (rx > 4L * n + 65536 || ry > 4L * n + 65536) return false;
    o.resize(n, n <= 65536);
    int32_t* ly = o.ly.data();
    for (int i = 0; i < n; i++) ly[i] = i;
    bool done = yxsorted;  // strictly increasing (y, x): no equal points either
    if (!done && ysorted) {
        // insertion sort of each row by x, while the moves stay within 8n
        long moves = 0;
        done = true;
        for (int s = 0; s < n && done;) {
            int e = s + 1;
            while (e < n && ky[e] == ky[s]) e++;
            for (int j = s + 1; j < e && moves <= 8L * n; j++) {
                const int32_t v = ly[j], kv = kx[v];
                int m = j;
                while (m > s && kx[ly[m - 1]] > kv) {
                    ly[m] = ly[m - 1];
                    m--;
                }
                moves += j - m;
                ly[m] = v;
            }
            if (moves > 8L * n) {
                done = false;
                break;
            }
            for (int j = s + 1; j < e; j++)
                if (kx[ly[j]] == kx[ly[j - 1]]) return false;
            s = e;
        }
    }
    if (!done) {
        tmp.resize(n);
        cnt.assign((size_t)rx + 1, 0);  // by x, then stably by y
        for (int i = 0; i < n; i++) cnt[kx[i] - lox + 1]++;
        for (long b = 0; b < rx; b++) cnt[b + 1] += cnt[b];
        for (int i = 0; i < n; i++) tmp[cnt[kx[i] - lox]++] = i;
        cnt.assign((size_t)ry + 1, 0);
        for (int i = 0; i < n; i++) cnt[ky[i] - loy + 1]++;
        for (long b = 0; b < ry; b++) cnt[b + 1] += cnt[b];
        for (int k = 0; k < n; k++) ly[cnt[ky[tmp[k]] - loy]++] = tmp[k];
        for (int k = 1; k < n; k++)
            if (kx[ly[k]] == kx[ly[k - 1]] && ky[ly[k]] == ky[ly[k - 1]]) return false;
    }
    // (x, y) order, ranks and entries: ly stably by x (a point's place in
    // that order is its x rank; cut() maps entries to points through ly)
    cnt.assign((size_t)rx + 1, 0);
    for (int i = 0; i < n; i++) cnt[kx[i] - lox + 1]++;
    for (long b = 0; b < rx; b++) cnt[b + 1] += cnt[b];
    int32_t* c = cnt.data();
    if (o.narrow) {
        uint32_t* px = o.qx.data();
        uint32_t* py = o.qy.data();
        for (int k = 0; k < n; k++) {
            const int32_t i = ly[k];
            const int32_t pos = c[kx[i] - lox]++;
            const uint32_t e = (uint32_t)pos << 16 | (uint32_t)k;
            px[pos] = e;
            py[k] = e;
        }
    } else {
        uint64_t* px = o.px.data();
        uint64_t* py = o.py.data();
        for (int k = 0; k < n; k++) {
            const int32_t i = ly[k];
            const int32_t pos = c[kx[i] - lox]++;
            const uint64_t e = (uint64_t)(uint32_t)pos << 32 | (uint32_t)k;
            px[pos] = e;
            py[k] = e;
        }
    }
    return true;
}

void elas_tri_order_sorted(const float* xy, int count, ElasTriOrder& o) {
    o.resize(count, false);
    for (int i = 0; i < count; i++) o.lx[i] = o.ly[i] = i;
    // (y, x) order: by y, the (x, y) order kept among equal y
    std::stable_sort(o.ly.begin(), o.ly.end(), [xy](int32_t a, int32_t b) { return xy[2 * a + 1] < xy[2 * b + 1]; });
    o.pack();
}

namespace {

// Split depth of a mesh of n points: the divide-and-conquer is cut into up to
// 2^depth concurrent subtrees of at least ELAS_TRI_LEAF_MIN points, at most
// g_split_max levels (elas_tri_set_split; -1 = ELAS_TRI_SPLIT_AUTO).
#define ELAS_TRI_LEAF_MIN 128
#define ELAS_TRI_SPLIT_AUTO 3
std::atomic<int> g_split_max{-1};

int split_depth(int n) {
    int mx = g_split_max.load(std::memory_order_relaxed);
    if (mx < 0) mx = ELAS_TRI_SPLIT_AUTO;
    const int leaf_min = g_split_max.load(std::memory_order_relaxed) < 0 ? ELAS_TRI_LEAF_MIN : 4;
    int d = 0;
    while (d < mx && (n >> (d + 1)) >= leaf_min) d++;
    return d;
}

void pool_parallel_for(void* ctx, int ntasks, void (*task)(void*, int), void* arg);
int pool_share(void* ctx);

// The depth a job gets when it starts: split_depth(), and under the automatic
// setting no more subtrees than the job's share of the threads (the jobs in
// flight split them; a batch of many meshes runs each one serially).
int job_depth(int n, void* pool) {
    int d = split_depth(n);
    if (g_split_max.load(std::memory_order_relaxed) < 0) {
        const int share = pool_share(pool);
        int cap = 0;
        while ((2 << cap) <= share) cap++;
        d = std::min(d, cap);
    }
    return d;
}

// One "zQB" Triangle call on copy k; with split set Triangle's setup and the
// vertex order are built at once, and the merge's subtrees run through the
// pool's helpers.  Corners are numbered by the caller's order.
struct TriCall {
    int k;
    const ElasTriJob* job;
    ElasTriOrder* order;
    void* setup;
    bool integer;
    int depth;
    void* pool;
};

// the cuts below the root down to the split depth, one half of the order
// per task
void tri_half_cut(void* arg, int t) {
    TriCall* c = (TriCall*)arg;
    const int n = c->job->n, h = n >> 1;
    c->order->cut(t ? h : 0, t ? n - h : h, 1, c->depth - 1);
}

// The vertex order and, when it was built here, the top `depth` levels of
// its cuts: the root's here, its halves' on two threads (task 0, which the
// calling thread takes first: the longer one); Triangle's setup on another
// thread
void tri_prepare(void* arg, int t) {
    TriCall* c = (TriCall*)arg;
    if (t == 1) {
        c->setup = kSetup[c->k](c->job->pts, c->job->n);
    } else {
        c->integer = elas_tri_order_integer(c->job->pts, c->job->n, *c->order);
        if (c->integer && c->depth > 0) {
            c->order->cut(0, c->job->n, 0, 1);
            if (c->depth > 1 && c->job->n > 3) pool_parallel_for(c->pool, 2, tri_half_cut, c);
        }
    }
}

// Brings a mesh of about n points' order scratch into this thread's cache
// (written, so the order's scattered stores find their lines here): run by
// the worker a mesh job is about to reach, while the device computes the
// support points.
void warm_order(int n) {
    OrderScratch& s = order_scratch();
    const bool narrow = n <= 65536;
    s.kxv.resize(n);
    s.kyv.resize(n);
    s.order.resize(n, narrow);
    memset(s.kxv.data(), 0, 4 * (size_t)n);
    memset(s.kyv.data(), 0, 4 * (size_t)n);
    memset(s.order.ly.data(), 0, 4 * (size_t)n);
    if (narrow) {
        for (std::vector<uint32_t>* v : {&s.order.qx, &s.order.qy, &s.order.ql, &s.order.qr})
            memset(v->data(), 0, 4 * (size_t)n);
    } else {
        for (std::vector<uint64_t>* v : {&s.order.px, &s.order.py, &s.order.tl, &s.order.tr})
            memset(v->data(), 0, 8 * (size_t)n);
    }
}

// Brings a block of `bytes` into this thread's cache and its own block list,
// from which the next pool of that size class this thread allocates comes
// (a subtree's triangle pool): run by each thread while the device computes
// the support points of a single-pair call.
void warm_block(size_t bytes) {
    void* b = elas_tri_block_alloc(bytes);
    if (!b) return;
    memset(b, 0, bytes);
    elas_tri_block_free(b);
}

int tri_call(int k, ElasTriJob& j, bool split, void* pool) {
    ElasTriOrder& order = order_scratch().order;
    const int depth = split ? job_depth(j.n, pool) : 0;
    TriCall c{k, &j, &order, nullptr, false, depth, pool};
    if (depth > 0) {
        pool_parallel_for(pool, 2, tri_prepare, &c);
    } else {
        tri_prepare(&c, 0);
        tri_prepare(&c, 1);
    }
    return kFinish[k](c.setup, &order, c.integer, depth, pool_parallel_for, pool, j.out, j.dst, j.cap, j.dst16);
}

int run_instance(int k, ElasTriJob& j, bool split, void* pool) {
    j.t_begin = now_s();
    if (j.n < 3) {  // Triangle would exit(1) (src/triangle.cpp:7652-7654)
        if (j.out) j.out->clear();
        j.t_end = now_s();
        return -1;
    }
    const int nt = tri_call(k, j, split, pool);
    j.t_end = now_s();
    return nt;
}

// Worker K owns Triangle copy K and runs whole jobs on it.  A job that splits
// its merge publishes each step's subtrees as a task set in one of a fixed
// number of slots; idle workers, and a caller waiting for its meshes, take
// tasks from the published sets before taking new jobs.  Tasks are claimed
// without a lock (an atomic counter per set); a slot's reader count keeps a
// set alive while a helper may still touch it.  Any thread may run a subtree
// of copy K: the merge only reads copy K's namespace-scope state, which copy
// K's owner wrote in triangleinit() before publishing.  A worker that just
// finished polls for work for up to ELAS_TRI_SPIN_S before sleeping: a core
// that went idle between calls runs Triangle up to ~20% slower for its first
// few hundred microseconds, which is the whole job at these sizes.
#define ELAS_TRI_SPIN_S 0.003
#define ELAS_TRI_SLOTS 32
class TriPool {
public:
    static TriPool& get() {
        static TriPool* pool = new TriPool();  // lives until process exit
        return *pool;
    }
    // remaining: jobs not finished, under m; left: the same count, lowered by
    // a worker as its last access of the batch (after releasing m), so a
    // waiter frees the batch only once left is 0
    struct Batch;
    struct Item {
        ElasTriJob* job;
        Batch* batch;
    };
    struct Batch {
        std::mutex m;
        std::condition_variable cv;
        int remaining = 0;
        std::atomic<int> left{0};
        Item* sticky[2] = {nullptr, nullptr};  // its jobs posted to a worker's box
        ~Batch() {
            delete sticky[0];
            delete sticky[1];
        }
    };
    // tasks of one step of a split merge, claimed through `next`; `done`
    // counts finished tasks
    struct TaskSet {
        void (*fn)(void*, int);
        void* arg;
        int n;
        std::atomic<int> next{0};
        std::atomic<int> done{0};
    };
    void run(ElasTriJob* jobs, int n, int max_workers) {
        if (n <= 0) return;
        if (n == 1 || max_workers <= 1) {
            // on the calling thread with copy 0; a single job still splits
            // its merge over the workers unless one worker is asked for
            std::lock_guard<std::mutex> g(serial_);
            if (max_workers > 1) ensure_workers();
            active_.fetch_add(n);
            for (int i = 0; i < n; i++) {
                jobs[i].result = run_instance(0, jobs[i], max_workers > 1, this);
                active_.fetch_sub(1);
            }
            return;
        }
        Batch b;
        enqueue(jobs, n, &b);
        {
            std::unique_lock<std::mutex> g(b.m);
            b.cv.wait(g, [&] { return b.remaining == 0; });
        }
        // a worker's last access of the batch is its decrement of `left`,
        // after it released m: b (on this stack) outlives it
        while (b.left.load(std::memory_order_acquire) > 0) _mm_pause();
    }
    // A batch of one or two jobs (the two meshes of one pair) goes to
    // workers 1 and 2 when their boxes are free, so each mesh of a run of
    // single-pair calls finds its thread's buffers where it left them; other
    // jobs go to the shared queue.
    // Asks workers 1 and 2, which take the meshes of single-pair calls, to
    // bring the order scratch of meshes of n0 and n1 points into their caches.
    // Every thread (the caller too) also brings into its cache a block of the
    // size a subtree's triangle pool takes.
    void prewarm(int n0, int n1) {
        // a subtree of an 8-way split: its pool's first block holds
        // 2 size + 8 triangles of 48 bytes ("zQB")
        const size_t leaf = 48 * (2 * (size_t)(std::max(n0, n1) / 8 + 1) + 8) + 16;
        {
            std::lock_guard<std::mutex> g(m_);
            ensure_workers();
            if (threads_.size() < 2) return;
            warm_[1].store(n0, std::memory_order_relaxed);
            warm_[2].store(n1, std::memory_order_relaxed);
            for (size_t k = 1; k <= threads_.size(); k++) warmb_[k].store((long)leaf, std::memory_order_relaxed);
        }
        cv_.notify_all();
        warm_block(leaf);
    }
    void enqueue(ElasTriJob* jobs, int n, Batch* b) {
        b->remaining = n;
        b->left.store(n, std::memory_order_relaxed);
        bool boxed = false;
        {
            std::lock_guard<std::mutex> g(m_);
            ensure_workers();
            int queued = 0;
            for (int i = 0; i < n; i++) {
                if (n <= 2 && (int)threads_.size() >= 2) {
                    b->sticky[i] = new Item{&jobs[i], b};
                    Item* expect = nullptr;
                    if (box_[i + 1].compare_exchange_strong(expect, b->sticky[i])) {
                        boxed = true;
                        continue;
                    }
                    delete b->sticky[i];
                    b->sticky[i] = nullptr;
                }
                q_.push_back({&jobs[i], b});
                queued++;
            }
            active_.fetch_add(n);
            pending_.fetch_add(queued, std::memory_order_release);
        }
        if (boxed) cv_.notify_all();
        else for (int i = 0; i < n; i++) cv_.notify_one();
    }
    // Runs the set's tasks here and on idle helpers; returns when all are done.
    void parallel_for(int ntasks, void (*fn)(void*, int), void* arg) {
        if (ntasks <= 0) return;
        TaskSet ts;
        ts.fn = fn;
        ts.arg = arg;
        ts.n = ntasks;
        int slot = -1;
        if (ntasks > 1) {
            for (int k = 0; k < ELAS_TRI_SLOTS && slot < 0; k++) {
                TaskSet* expect = nullptr;
                if (slots_[k].compare_exchange_strong(expect, &ts)) slot = k;
            }
            if (slot >= 0) {  // (no free slot: the caller runs every task)
                helping_.fetch_add(1);
                { std::lock_guard<std::mutex> g(m_); }  // a worker between its check and its sleep sees the set
                for (int i = 1; i < ntasks; i++) cv_.notify_one();
            }
        }
        for (;;) {
            const int i = ts.next.fetch_add(1);
            if (i >= ntasks) break;
            fn(arg, i);
            ts.done.fetch_add(1, std::memory_order_release);
        }
        if (slot >= 0) {
            // unpublish, then wait until no helper can still hold the set
            slots_[slot].store(nullptr);
            helping_.fetch_sub(1);
            while (readers_[slot].load() != 0) _mm_pause();
        }
        while (ts.done.load(std::memory_order_acquire) < ntasks) _mm_pause();
    }
    // Runs one task of a published set on this thread; false if there is none.
    // The slot's reader count covers only the claim: once a task is claimed,
    // its set lives until its publisher has counted it done, and a task may
    // itself publish a set (which may land in this very slot, whose readers
    // its publisher waits out).
    bool help_one() {
        if (helping_.load() == 0) return false;
        for (int k = 0; k < ELAS_TRI_SLOTS; k++) {
            if (!slots_[k].load(std::memory_order_relaxed)) continue;
            readers_[k].fetch_add(1);
            TaskSet* ts = slots_[k].load();
            int i = -1;
            if (ts) {
                i = ts->next.fetch_add(1);
                if (i >= ts->n) i = -1;
            }
            readers_[k].fetch_sub(1);
            if (i >= 0) {
                ts->fn(ts->arg, i);
                ts->done.fetch_add(1, std::memory_order_release);
                return true;
            }
        }
        return false;
    }

private:
    // One worker per CPU this process may run on, less one for the caller
    // (at most ELAS_TRI_INSTANCES - 1: copy 0 is the caller's): a worker
    // that waits for a CPU while holding a claimed subtree stalls the whole
    // merge, and idle workers poll before they sleep.
    void ensure_workers() {
        if (!threads_.empty()) return;
        int cpus = (int)std::thread::hardware_concurrency();
        cpu_set_t set;
        if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = CPU_COUNT(&set);
        const int nw = std::max(1, std::min(ELAS_TRI_INSTANCES - 1, cpus - 1));
        for (int k = 1; k <= nw; k++) threads_.emplace_back([this, k] { loop(k); });
    }
    void loop(int k) {
        double last = 0.0;  // end of this worker's last job or task
        for (;;) {
            if (help_one()) {  // subtrees of a running job first
                last = now_s();
                continue;
            }
            if (const int w = warm_[k].exchange(0, std::memory_order_relaxed)) {
                warm_order(w);
                last = now_s();
                continue;
            }
            if (const long wb = warmb_[k].exchange(0, std::memory_order_relaxed)) {
                warm_block((size_t)wb);
                last = now_s();
                continue;
            }
            Item it;
            if (Item* bx = box_[k].exchange(nullptr, std::memory_order_acquire)) {
                it = *bx;  // the Item itself is freed with its batch
            } else {
                if (pending_.load(std::memory_order_acquire) == 0) {
                    const double idle = now_s() - last;
                    if (idle < ELAS_TRI_SPIN_S) {
                        // poll; after the first microseconds, give the CPU
                        // to any other thread that wants it meanwhile
                        if (idle < 20e-6) _mm_pause();
                        else std::this_thread::yield();
                        continue;
                    }
                    std::unique_lock<std::mutex> g(m_);
                    cv_.wait(g, [&] {
                        return !q_.empty() || helping_.load() != 0 || box_[k].load() != nullptr || warm_[k].load() != 0 ||
                               warmb_[k].load() != 0;
                    });
                    last = now_s();  // woken: poll for a while before sleeping again
                    continue;
                }
                std::lock_guard<std::mutex> g(m_);
                if (q_.empty()) continue;
                it = q_.front();
                q_.pop_front();
                pending_.fetch_sub(1, std::memory_order_relaxed);
            }
            it.job->result = run_instance(k, *it.job, true, this);
            active_.fetch_sub(1);
            {
                std::lock_guard<std::mutex> g(it.batch->m);
                if (--it.batch->remaining == 0) it.batch->cv.notify_all();
            }
            it.batch->left.fetch_sub(1, std::memory_order_release);
            last = now_s();
        }
    }
public:
    // threads that may run a job's subtrees: the workers and the caller
    int share() const {
        const int act = std::max(1, active_.load(std::memory_order_relaxed));
        return std::max(1, ((int)threads_.size() + 1) / act);
    }

private:
    std::atomic<int> pending_{0};  // queued jobs, read without the lock while polling
    std::atomic<int> active_{0};   // jobs queued or running
    std::atomic<int> helping_{0};  // published task sets
    std::atomic<TaskSet*> slots_[ELAS_TRI_SLOTS] = {};
    std::atomic<Item*> box_[ELAS_TRI_INSTANCES] = {};  // a job for worker k
    std::atomic<int> warm_[ELAS_TRI_INSTANCES] = {};   // points of a mesh worker k is about to get
    std::atomic<long> warmb_[ELAS_TRI_INSTANCES] = {};  // bytes of a pool block worker k warms
    std::atomic<int> readers_[ELAS_TRI_SLOTS] = {};
    std::mutex m_, serial_;
    std::condition_variable cv_;
    std::deque<Item> q_;
    std::vector<std::thread> threads_;
};

int pool_share(void* ctx) { return static_cast<TriPool*>(ctx)->share(); }

void pool_parallel_for(void* ctx, int ntasks, void (*task)(void*, int), void* arg) {
    static_cast<TriPool*>(ctx)->parallel_for(ntasks, task, arg);
}

}  // namespace

void elas_triangulate_jobs(ElasTriJob* jobs, int n, int max_workers) {
    TriPool::get().run(jobs, n, max_workers);
}

void elas_tri_prewarm(int n0, int n1) {
    if (n0 >= 3 && n1 >= 3) TriPool::get().prewarm(n0, n1);
}

void elas_tri_set_split(int max_depth) { g_split_max.store(max_depth < 0 ? -1 : std::min(max_depth, 12)); }

int elas_tri_split_depth(int n) { return split_depth(n); }

struct ElasTriTicket {
    TriPool::Batch batch;
};

ElasTriTicket* elas_tri_submit(ElasTriJob* jobs, int n) {
    ElasTriTicket* t = new ElasTriTicket();
    if (n <= 0) {
        t->batch.remaining = 0;
        return t;
    }
    TriPool::get().enqueue(jobs, n, &t->batch);
    return t;
}

bool elas_tri_done(ElasTriTicket* t) {
    std::lock_guard<std::mutex> g(t->batch.m);
    return t->batch.remaining == 0;
}

// Polls for up to ELAS_TRI_SPIN_S before sleeping, running subtrees of the
// jobs' split merges meanwhile: a synchronous call waits here for meshes that
// take a fraction of a millisecond, and a sleeping thread's wake-up would add
// tens of microseconds to it.
void elas_tri_wait(ElasTriTicket* t) {
    const double t0 = now_s();
    TriPool& pool = TriPool::get();
    while (t->batch.left.load(std::memory_order_acquire) > 0 && now_s() - t0 < ELAS_TRI_SPIN_S)
        if (!pool.help_one()) _mm_pause();
    {
        std::unique_lock<std::mutex> g(t->batch.m);
        t->batch.cv.wait(g, [&] { return t->batch.remaining == 0; });
    }
    // the ticket may be freed once no worker can touch it: a worker's last
    // access is its decrement of `left`, after it released m
    while (t->batch.left.load(std::memory_order_acquire) > 0) _mm_pause();
}

void elas_tri_release(ElasTriTicket* t) {
    if (!t) return;
    elas_tri_wait(t);
    delete t;
}
