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
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
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

#define ELAS_TRI_DECL(NS)                                                                          \
    namespace NS {                                                                                 \
    int elas_triangulate_arranged(const float* pts, int n, const int* arr, int** trilist);         \
    int elas_triangulate_dedup(const float* pts, int n, bool (*arrange)(const float*, int, int*), \
                               int** trilist);                                                     \
    int elas_triangulate_split(const float* pts, int n, const int* arr, int depth,                 \
                               ElasParallelFor par, void* parctx, std::vector<int32_t>* out);      \
    int elas_triangulate_dedup_split(const float* pts, int n,                                      \
                                     bool (*arrange)(const float*, int, int*), int depth,          \
                                     ElasParallelFor par, void* parctx, std::vector<int32_t>* out); \
    }
ELAS_TRI_DECL(elas_tri0) ELAS_TRI_DECL(elas_tri1) ELAS_TRI_DECL(elas_tri2) ELAS_TRI_DECL(elas_tri3)
ELAS_TRI_DECL(elas_tri4) ELAS_TRI_DECL(elas_tri5) ELAS_TRI_DECL(elas_tri6) ELAS_TRI_DECL(elas_tri7)
ELAS_TRI_DECL(elas_tri8) ELAS_TRI_DECL(elas_tri9) ELAS_TRI_DECL(elas_tri10) ELAS_TRI_DECL(elas_tri11)
ELAS_TRI_DECL(elas_tri12) ELAS_TRI_DECL(elas_tri13) ELAS_TRI_DECL(elas_tri14) ELAS_TRI_DECL(elas_tri15)

namespace {

typedef int (*DedupFn)(const float*, int, bool (*)(const float*, int, int*), int**);
const DedupFn kDedup[ELAS_TRI_INSTANCES] = {
    elas_tri0::elas_triangulate_dedup, elas_tri1::elas_triangulate_dedup,
    elas_tri2::elas_triangulate_dedup, elas_tri3::elas_triangulate_dedup,
    elas_tri4::elas_triangulate_dedup, elas_tri5::elas_triangulate_dedup,
    elas_tri6::elas_triangulate_dedup, elas_tri7::elas_triangulate_dedup,
    elas_tri8::elas_triangulate_dedup, elas_tri9::elas_triangulate_dedup,
    elas_tri10::elas_triangulate_dedup, elas_tri11::elas_triangulate_dedup,
    elas_tri12::elas_triangulate_dedup, elas_tri13::elas_triangulate_dedup,
    elas_tri14::elas_triangulate_dedup, elas_tri15::elas_triangulate_dedup};

typedef int (*ArrFn)(const float*, int, const int*, int**);
const ArrFn kArr[ELAS_TRI_INSTANCES] = {
    elas_tri0::elas_triangulate_arranged, elas_tri1::elas_triangulate_arranged,
    elas_tri2::elas_triangulate_arranged, elas_tri3::elas_triangulate_arranged,
    elas_tri4::elas_triangulate_arranged, elas_tri5::elas_triangulate_arranged,
    elas_tri6::elas_triangulate_arranged, elas_tri7::elas_triangulate_arranged,
    elas_tri8::elas_triangulate_arranged, elas_tri9::elas_triangulate_arranged,
    elas_tri10::elas_triangulate_arranged, elas_tri11::elas_triangulate_arranged,
    elas_tri12::elas_triangulate_arranged, elas_tri13::elas_triangulate_arranged,
    elas_tri14::elas_triangulate_arranged, elas_tri15::elas_triangulate_arranged};

typedef int (*SplitFn)(const float*, int, const int*, int, ElasParallelFor, void*, std::vector<int32_t>*);
const SplitFn kSplit[ELAS_TRI_INSTANCES] = {
    elas_tri0::elas_triangulate_split, elas_tri1::elas_triangulate_split,
    elas_tri2::elas_triangulate_split, elas_tri3::elas_triangulate_split,
    elas_tri4::elas_triangulate_split, elas_tri5::elas_triangulate_split,
    elas_tri6::elas_triangulate_split, elas_tri7::elas_triangulate_split,
    elas_tri8::elas_triangulate_split, elas_tri9::elas_triangulate_split,
    elas_tri10::elas_triangulate_split, elas_tri11::elas_triangulate_split,
    elas_tri12::elas_triangulate_split, elas_tri13::elas_triangulate_split,
    elas_tri14::elas_triangulate_split, elas_tri15::elas_triangulate_split};

typedef int (*DedupSplitFn)(const float*, int, bool (*)(const float*, int, int*), int, ElasParallelFor, void*,
                            std::vector<int32_t>*);
const DedupSplitFn kDedupSplit[ELAS_TRI_INSTANCES] = {
    elas_tri0::elas_triangulate_dedup_split, elas_tri1::elas_triangulate_dedup_split,
    elas_tri2::elas_triangulate_dedup_split, elas_tri3::elas_triangulate_dedup_split,
    elas_tri4::elas_triangulate_dedup_split, elas_tri5::elas_triangulate_dedup_split,
    elas_tri6::elas_triangulate_dedup_split, elas_tri7::elas_triangulate_dedup_split,
    elas_tri8::elas_triangulate_dedup_split, elas_tri9::elas_triangulate_dedup_split,
    elas_tri10::elas_triangulate_dedup_split, elas_tri11::elas_triangulate_dedup_split,
    elas_tri12::elas_triangulate_dedup_split, elas_tri13::elas_triangulate_dedup_split,
    elas_tri14::elas_triangulate_dedup_split, elas_tri15::elas_triangulate_dedup_split};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

namespace {

// Stable counting sort of the indices in `in` by the integer keys key[]
// (already offset to 0..range-1).
void counting_sort(const int32_t* key, int n, long range, const int32_t* in, int32_t* out,
                   std::vector<int32_t>& cnt) {
    cnt.assign((size_t)range + 1, 0);
    for (int i = 0; i < n; i++) cnt[key[in[i]] + 1]++;
    for (long b = 0; b < range; b++) cnt[b + 1] += cnt[b];
    for (int i = 0; i < n; i++) out[cnt[key[in[i]]]++] = in[i];
}

// Alternating cuts over a node [lo, lo + size): lx and ly hold the node's
// points in (x, y) and (y, x) order.  A cut along `axis` keeps the size / 2
// smallest points of that order on its left (as vertexmedian() does,
// src/triangle.cpp:5517-5574) and splits the other list stably by side, so
// both children keep both orders.  Nodes of 2 or 3 points are leaves sorted
// by (x, y) (alternateaxes(), src/triangle.cpp:5586-5605).
struct Cuts {
    int32_t *lx, *ly, *tmp, *out;
    uint8_t* side;
    void node(int lo, int size, int axis) {
        if (size <= 3) {
            for (int k = 0; k < size; k++) out[lo + k] = lx[lo + k];
            return;
        }
        const int divider = size >> 1;
        int32_t* cut = axis ? ly : lx;
        int32_t* other = axis ? lx : ly;
        for (int k = 0; k < divider; k++) side[cut[lo + k]] = 0;
        for (int k = divider; k < size; k++) side[cut[lo + k]] = 1;
        // branch-free stable split: the side of a point is not predictable
        int a = 0, b = divider;
        for (int k = 0; k < size; k++) {
            const int32_t i = other[lo + k];
            const int r = side[i];
            tmp[r ? b : a] = i;
            a += 1 - r;
            b += r;
        }
        memcpy(other + lo, tmp, sizeof(int32_t) * (size_t)size);
        node(lo, divider, 1 - axis);
        node(lo + divider, size - divider, 1 - axis);
    }
};

}  // namespace

// Triangle's divide-and-conquer order (divconqdelaunay(), src/triangle.cpp:
// 6164-6230): the points sorted by (x, y), halved, and each half cut
// alternately along y and x down to 2- and 3-point leaves.  For distinct
// points that order is unique, so it can be built here in O(n log n) from two
// presorted lists instead of by Triangle's randomised quicksort and
// quickselect, which take about half of a support-point mesh's time.  Support
// points have integer coordinates, so both lists come from counting sorts.
// False when a coordinate is not a bounded integer or two points coincide
// (which duplicate Triangle keeps depends on its own sort); the caller then
// lets Triangle sort and drop duplicates itself and builds only the cuts.
bool elas_tri_arrangement(const float* p, int n, std::vector<int32_t>& arr) {
    if (n < 3) return false;
    thread_local std::vector<int32_t> cnt, a, b, lx, ly, tmp, key[2];
    thread_local std::vector<uint8_t> side;
    int lo[2] = {0, 0}, hi[2] = {0, 0};
    for (int c = 0; c < 2; c++) {
        key[c].resize(n);
        int32_t* kc = key[c].data();
        lo[c] = hi[c] = 0;
        for (int i = 0; i < n; i++) {
            const float v = p[2 * i + c];
            if (!(std::fabs(v) < 16777216.0f)) return false;
            const int iv = (int)v;
            if ((float)iv != v) return false;
            kc[i] = iv;
            if (i == 0 || iv < lo[c]) lo[c] = iv;
            if (i == 0 || iv > hi[c]) hi[c] = iv;
        }
    }
    const long rx = (long)hi[0] - lo[0] + 1, ry = (long)hi[1] - lo[1] + 1;
    if (rx > 4L * n + 65536 || ry > 4L * n + 65536) return false;
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < n; i++) key[c][i] -= lo[c];
    a.resize(n);
    b.resize(n);
    lx.resize(n);
    ly.resize(n);
    for (int i = 0; i < n; i++) a[i] = i;
    counting_sort(key[1].data(), n, ry, a.data(), b.data(), cnt);   // by y
    counting_sort(key[0].data(), n, rx, b.data(), lx.data(), cnt);  // then x: (x, y)
    counting_sort(key[0].data(), n, rx, a.data(), b.data(), cnt);   // by x
    counting_sort(key[1].data(), n, ry, b.data(), ly.data(), cnt);  // then y: (y, x)
    for (int k = 1; k < n; k++) {
        const int32_t i = lx[k - 1], j = lx[k];
        if (key[0][i] == key[0][j] && key[1][i] == key[1][j]) return false;
    }
    tmp.resize(n);
    side.resize(n);
    arr.resize(n);
    Cuts cuts{lx.data(), ly.data(), tmp.data(), arr.data(), side.data()};
    cuts.node(0, n, 0);
    return true;
}

namespace {

// One "zQB" Triangle call on copy k.  Corners are numbered by the caller's
// order either way.
// The alternating cuts of distinct points given in (x, y) order: positions
// into xy.  Used after Triangle's own sort and duplicate removal, for sets the
// counting sorts do not take (duplicates, non-integer coordinates).
bool arrange_sorted(const float* xy, int count, int* arr) {
    thread_local std::vector<int32_t> lx, ly, tmp;
    thread_local std::vector<uint8_t> side;
    lx.resize(count);
    ly.resize(count);
    tmp.resize(count);
    side.resize(count);
    for (int i = 0; i < count; i++) lx[i] = ly[i] = i;
    // (y, x) order: by y, the (x, y) order kept among equal y
    std::stable_sort(ly.begin(), ly.end(), [xy](int32_t a, int32_t b) { return xy[2 * a + 1] < xy[2 * b + 1]; });
    Cuts cuts{lx.data(), ly.data(), tmp.data(), arr, side.data()};
    cuts.node(0, count, 0);
    return true;
}

// Split depth of a mesh of n points: the divide-and-conquer is cut into up to
// 2^depth concurrent subtrees of at least ELAS_TRI_LEAF_MIN points, at most
// g_split_max levels (elas_tri_set_split; -1 = ELAS_TRI_SPLIT_AUTO).
#define ELAS_TRI_LEAF_MIN 384
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

// One "zQB" Triangle call on copy k; with split > 0 the merge's subtrees run
// through the pool's helpers.  Corners are numbered by the caller's order.
int tri_call(int k, const float* pts, int n, std::vector<int32_t>& out, bool split, void* pool) {
    thread_local std::vector<int32_t> arr;
    const int depth = split ? split_depth(n) : 0;
    if (depth > 0) {
        const int nt = elas_tri_arrangement(pts, n, arr)
                           ? kSplit[k](pts, n, arr.data(), depth, pool_parallel_for, pool, &out)
                           : kDedupSplit[k](pts, n, arrange_sorted, depth, pool_parallel_for, pool, &out);
        if (nt < 0) out.clear();
        return nt;
    }
    int* tl = nullptr;
    const int nt = elas_tri_arrangement(pts, n, arr) ? kArr[k](pts, n, arr.data(), &tl)
                                                     : kDedup[k](pts, n, arrange_sorted, &tl);
    if (nt < 0) {  // a single distinct point
        out.clear();
        return -1;
    }
    out.assign(tl, tl + 3 * (size_t)nt);
    free(tl);
    return nt;
}

int run_instance(int k, ElasTriJob& j, bool split, void* pool) {
    j.t_begin = now_s();
    if (j.n < 3) {  // Triangle would exit(1) (src/triangle.cpp:7652-7654)
        j.out->clear();
        j.t_end = now_s();
        return -1;
    }
    const int nt = tri_call(k, j.pts, j.n, *j.out, split, pool);
    j.t_end = now_s();
    return nt;
}

// Worker K owns Triangle copy K and runs whole jobs on it.  A job that splits
// its merge publishes the subtrees as a task set; idle workers, and a caller
// waiting for its meshes, take tasks from the sets before taking new jobs.
// Any thread may run a subtree of copy K: the merge only reads copy K's
// namespace-scope state, which copy K's owner wrote in triangleinit() before
// publishing the set.  A worker that just finished polls for work for up to
// ELAS_TRI_SPIN_S before sleeping: a core that went idle between calls runs
// Triangle up to ~20% slower for its first few hundred microseconds, which is
// the whole job at these sizes.
#define ELAS_TRI_SPIN_S 0.003
class TriPool {
public:
    static TriPool& get() {
        static TriPool* pool = new TriPool();  // lives until process exit
        return *pool;
    }
    // remaining: jobs not finished, under m; left: the same count, lowered by
    // a worker as its last access of the batch, so a waiter that polls it and
    // then takes m once may free the batch
    struct Batch {
        std::mutex m;
        std::condition_variable cv;
        int remaining = 0;
        std::atomic<int> left{0};
    };
    // tasks of one split merge; claimed through `next` (by helpers under m_),
    // `done` counts finished tasks and is each helper's last access
    struct TaskSet {
        void (*fn)(void*, int);
        void* arg;
        int n;
        std::atomic<int> next{0};
        std::atomic<int> done{0};
        bool listed = false;  // in help_, under m_
    };
    void run(ElasTriJob* jobs, int n, int max_workers) {
        if (n <= 0) return;
        if (n == 1 || max_workers <= 1) {
            // on the calling thread with copy 0; a single job still splits
            // its merge over the workers unless one worker is asked for
            std::lock_guard<std::mutex> g(serial_);
            if (max_workers > 1) ensure_workers();
            for (int i = 0; i < n; i++) jobs[i].result = run_instance(0, jobs[i], max_workers > 1, this);
            return;
        }
        Batch b;
        enqueue(jobs, n, &b);
        std::unique_lock<std::mutex> g(b.m);
        b.cv.wait(g, [&] { return b.remaining == 0; });
    }
    void enqueue(ElasTriJob* jobs, int n, Batch* b) {
        b->remaining = n;
        b->left.store(n, std::memory_order_relaxed);
        {
            std::lock_guard<std::mutex> g(m_);
            ensure_workers();
            for (int i = 0; i < n; i++) q_.push_back({&jobs[i], b});
            pending_.fetch_add(n, std::memory_order_release);
        }
        for (int i = 0; i < n; i++) cv_.notify_one();
    }
    // Runs the set's tasks here and on idle helpers; returns when all are done.
    void parallel_for(int ntasks, void (*fn)(void*, int), void* arg) {
        if (ntasks <= 0) return;
        TaskSet ts;
        ts.fn = fn;
        ts.arg = arg;
        ts.n = ntasks;
        if (ntasks > 1) {
            {
                std::lock_guard<std::mutex> g(m_);
                ts.listed = true;
                help_.push_back(&ts);
                helping_.fetch_add(1, std::memory_order_release);
            }
            for (int i = 1; i < ntasks; i++) cv_.notify_one();
        }
        for (;;) {
            const int i = ts.next.fetch_add(1, std::memory_order_relaxed);
            if (i >= ntasks) break;
            fn(arg, i);
            ts.done.fetch_add(1, std::memory_order_release);
        }
        if (ntasks > 1) {
            std::lock_guard<std::mutex> g(m_);
            unlist(&ts);
        }
        while (ts.done.load(std::memory_order_acquire) < ntasks) _mm_pause();
    }
    // Runs one task of a listed set on this thread; false if there is none.
    bool help_one() {
        if (helping_.load(std::memory_order_acquire) == 0) return false;
        TaskSet* ts = nullptr;
        int i = 0;
        {
            std::lock_guard<std::mutex> g(m_);
            if (!claim(&ts, &i)) return false;
        }
        ts->fn(ts->arg, i);
        ts->done.fetch_add(1, std::memory_order_release);
        return true;
    }

private:
    struct Item {
        ElasTriJob* job;
        Batch* batch;
    };
    void unlist(TaskSet* ts) {
        if (!ts->listed) return;
        ts->listed = false;
        help_.erase(std::find(help_.begin(), help_.end(), ts));
        helping_.fetch_sub(1, std::memory_order_release);
    }
    // Claims a task of the first listed set that has one (m_ held); exhausted
    // sets are unlisted on the way.
    bool claim(TaskSet** out, int* idx) {
        while (!help_.empty()) {
            TaskSet* ts = help_.front();
            const int i = ts->next.fetch_add(1, std::memory_order_relaxed);
            if (i < ts->n) {
                *out = ts;
                *idx = i;
                return true;
            }
            unlist(ts);
        }
        return false;
    }
    void ensure_workers() {
        if (!threads_.empty()) return;
        for (int k = 1; k < ELAS_TRI_INSTANCES; k++) threads_.emplace_back([this, k] { loop(k); });
    }
    void loop(int k) {
        double last = 0.0;  // end of this worker's last job or task
        for (;;) {
            Item it;
            {
                while (pending_.load(std::memory_order_acquire) == 0 && helping_.load(std::memory_order_acquire) == 0 &&
                       now_s() - last < ELAS_TRI_SPIN_S)
                    _mm_pause();
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return !q_.empty() || !help_.empty(); });
                TaskSet* ts = nullptr;
                int i = 0;
                if (claim(&ts, &i)) {  // subtrees of a running job first
                    g.unlock();
                    ts->fn(ts->arg, i);
                    ts->done.fetch_add(1, std::memory_order_release);
                    last = now_s();
                    continue;
                }
                if (q_.empty()) continue;
                it = q_.front();
                q_.pop_front();
                pending_.fetch_sub(1, std::memory_order_relaxed);
            }
            it.job->result = run_instance(k, *it.job, true, this);
            {
                std::lock_guard<std::mutex> g(it.batch->m);
                if (--it.batch->remaining == 0) it.batch->cv.notify_all();
            }
            it.batch->left.fetch_sub(1, std::memory_order_release);
            last = now_s();
        }
    }
    std::atomic<int> pending_{0};  // queued jobs, read without the lock while polling
    std::atomic<int> helping_{0};  // listed task sets, read without the lock while polling
    std::mutex m_, serial_;
    std::condition_variable cv_;
    std::deque<Item> q_;
    std::vector<TaskSet*> help_;
    std::vector<std::thread> threads_;
};

void pool_parallel_for(void* ctx, int ntasks, void (*task)(void*, int), void* arg) {
    static_cast<TriPool*>(ctx)->parallel_for(ntasks, task, arg);
}

}  // namespace

void elas_triangulate_jobs(ElasTriJob* jobs, int n, int max_workers) {
    TriPool::get().run(jobs, n, max_workers);
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
    std::unique_lock<std::mutex> g(t->batch.m);
    t->batch.cv.wait(g, [&] { return t->batch.remaining == 0; });
}

void elas_tri_release(ElasTriTicket* t) {
    if (!t) return;
    elas_tri_wait(t);
    delete t;
}
