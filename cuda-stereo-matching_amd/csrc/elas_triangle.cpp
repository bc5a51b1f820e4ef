// elas_triangle.cpp -- host Delaunay bridge over the reference's Triangle.
//
// The reference triangulates on the host with TRILIB::triangulate("zQB")
// (src/Elas.cu:282-328).  Triangle keeps its state in namespace-scope
// globals (src/triangle.cpp:545-554, the RNG reset at :4034), so one instance
// is not reentrant.  The build compiles the reference's triangle.cpp
// unmodified ELAS_TRI_INSTANCES times with -DTRILIB=elas_triK: each copy gets
// its own namespace and globals, and worker thread K of the pool below uses
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

namespace {

// "zQB": zero-based indices, quiet, no boundary markers (Elas.cu:315-317).
#define ELAS_TRI_RUN(NS)                                                        \
    int run_##NS(const float* pts, int n, std::vector<int32_t>& out) {          \
        NS::triangulateio in, o;                                                \
        memset(&in, 0, sizeof(in));                                             \
        memset(&o, 0, sizeof(o));                                               \
        float* copy = (float*)malloc(sizeof(float) * 2 * (size_t)n);            \
        if (!copy) return -1;                                                   \
        memcpy(copy, pts, sizeof(float) * 2 * (size_t)n);                       \
        in.numberofpoints = n;                                                  \
        in.pointlist = copy;                                                    \
        char sw[] = "zQB";                                                      \
        NS::triangulate(sw, &in, &o, nullptr);                                  \
        const int nt = o.numberoftriangles;                                     \
        out.assign(o.trianglelist, o.trianglelist + 3 * (size_t)nt);            \
        free(copy);                                                             \
        free(o.pointlist);                                                      \
        free(o.trianglelist);                                                   \
        return nt;                                                              \
    }

ELAS_TRI_RUN(elas_tri0) ELAS_TRI_RUN(elas_tri1) ELAS_TRI_RUN(elas_tri2) ELAS_TRI_RUN(elas_tri3)
ELAS_TRI_RUN(elas_tri4) ELAS_TRI_RUN(elas_tri5) ELAS_TRI_RUN(elas_tri6) ELAS_TRI_RUN(elas_tri7)
ELAS_TRI_RUN(elas_tri8) ELAS_TRI_RUN(elas_tri9) ELAS_TRI_RUN(elas_tri10) ELAS_TRI_RUN(elas_tri11)
ELAS_TRI_RUN(elas_tri12) ELAS_TRI_RUN(elas_tri13) ELAS_TRI_RUN(elas_tri14) ELAS_TRI_RUN(elas_tri15)

typedef int (*TriFn)(const float*, int, std::vector<int32_t>&);
const TriFn kTri[ELAS_TRI_INSTANCES] = {
    run_elas_tri0, run_elas_tri1, run_elas_tri2, run_elas_tri3, run_elas_tri4, run_elas_tri5,
    run_elas_tri6, run_elas_tri7, run_elas_tri8, run_elas_tri9, run_elas_tri10, run_elas_tri11,
    run_elas_tri12, run_elas_tri13, run_elas_tri14, run_elas_tri15};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// Triangle's mesh of a point set without duplicate points does not depend on
// the order of the points: divconqdelaunay() first sorts its vertex pointers
// by (x, y) (src/triangle.cpp:6176-6198) and from there on works on that
// array and the mesh, and the output numbers each corner by its input
// position (writenodes, src/triangle.cpp:7708-7768).  Its quicksort (src/triangle.cpp:
// 5450-5505) runs about a quarter faster on sorted input, so a job hands
// Triangle its points in (x, y) order and maps the corners back.  Support
// points have integer coordinates, and those of one image arrive in row
// order, so a counting sort on x that keeps the input order among equal x
// yields (x, y) order.  The order is used only when the result is strictly
// increasing in (x, y): a set with duplicates keeps the caller's order, as
// which duplicate Triangle keeps depends on it.
bool elas_tri_sorted_order(const float* p, int n, std::vector<int32_t>& order, std::vector<float>& sorted) {
    if (n < 2) return false;
    float xmin = p[0], xmax = p[0];
    for (int i = 0; i < n; i++) {
        const float x = p[2 * i];
        if (!(std::fabs(x) < 16777216.0f) || x != std::floor(x)) return false;
        xmin = std::min(xmin, x);
        xmax = std::max(xmax, x);
    }
    const long range = (long)xmax - (long)xmin + 1;
    if (range > 4L * n + 65536) return false;
    thread_local std::vector<int32_t> cnt;
    cnt.assign((size_t)range + 1, 0);
    const int x0 = (int)xmin;
    for (int i = 0; i < n; i++) cnt[(int)p[2 * i] - x0 + 1]++;
    for (long b = 0; b < range; b++) cnt[b + 1] += cnt[b];
    order.resize(n);
    for (int i = 0; i < n; i++) order[cnt[(int)p[2 * i] - x0]++] = i;
    sorted.resize(2 * (size_t)n);
    for (int k = 0; k < n; k++) {
        sorted[2 * k] = p[2 * order[k]];
        sorted[2 * k + 1] = p[2 * order[k] + 1];
    }
    for (int k = 1; k < n; k++) {
        const float xa = sorted[2 * k - 2], ya = sorted[2 * k - 1], xb = sorted[2 * k], yb = sorted[2 * k + 1];
        if (!(xa < xb || (xa == xb && ya < yb))) return false;
    }
    return true;
}

namespace {

// One "zQB" Triangle call on copy k: points handed over in (x, y) order where
// that is safe, corners numbered by the caller's order.
int tri_call(int k, const float* pts, int n, std::vector<int32_t>& out) {
    thread_local std::vector<int32_t> order;
    thread_local std::vector<float> sorted;
    if (!elas_tri_sorted_order(pts, n, order, sorted)) return kTri[k](pts, n, out);
    const int nt = kTri[k](sorted.data(), n, out);
    for (auto& c : out) c = order[c];
    return nt;
}

int run_instance(int k, ElasTriJob& j) {
    j.t_begin = now_s();
    if (j.n < 3) {  // Triangle would exit(1) (src/triangle.cpp:7652-7654)
        j.out->clear();
        j.t_end = now_s();
        return -1;
    }
    const int nt = tri_call(k, j.pts, j.n, *j.out);
    j.t_end = now_s();
    return nt;
}

// Worker K owns Triangle copy K.  Jobs of one submit() are spread over the
// workers; submit() returns when all of them are done.  A worker that just
// finished a job polls the queue for up to ELAS_TRI_SPIN_S before sleeping: a
// core that went idle between calls runs Triangle up to ~20% slower for its
// first few hundred microseconds, which is the whole job at these sizes.
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
    void run(ElasTriJob* jobs, int n, int max_workers) {
        if (n <= 0) return;
        if (n == 1 || max_workers <= 1) {
            std::lock_guard<std::mutex> g(serial_);
            for (int i = 0; i < n; i++) jobs[i].result = run_instance(0, jobs[i]);
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

private:
    struct Item {
        ElasTriJob* job;
        Batch* batch;
    };
    void ensure_workers() {
        if (!threads_.empty()) return;
        for (int k = 1; k < ELAS_TRI_INSTANCES; k++) threads_.emplace_back([this, k] { loop(k); });
    }
    void loop(int k) {
        double last = 0.0;  // end of this worker's last job
        for (;;) {
            Item it;
            {
                while (pending_.load(std::memory_order_acquire) == 0 && now_s() - last < ELAS_TRI_SPIN_S) _mm_pause();
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return !q_.empty(); });
                it = q_.front();
                q_.pop_front();
                pending_.fetch_sub(1, std::memory_order_relaxed);
            }
            it.job->result = run_instance(k, *it.job);
            {
                std::lock_guard<std::mutex> g(it.batch->m);
                if (--it.batch->remaining == 0) it.batch->cv.notify_all();
            }
            it.batch->left.fetch_sub(1, std::memory_order_release);
            last = now_s();
        }
    }
    std::atomic<int> pending_{0};  // queued jobs, read without the lock while polling
    std::mutex m_, serial_;
    std::condition_variable cv_;
    std::deque<Item> q_;
    std::vector<std::thread> threads_;
};

}  // namespace

void elas_triangulate_jobs(ElasTriJob* jobs, int n, int max_workers) {
    TriPool::get().run(jobs, n, max_workers);
}

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

// Polls for up to ELAS_TRI_SPIN_S before sleeping: a synchronous call waits
// here for meshes that take a fraction of a millisecond, and a sleeping
// thread's wake-up would add tens of microseconds to it.
void elas_tri_wait(ElasTriTicket* t) {
    const double t0 = now_s();
    while (t->batch.left.load(std::memory_order_acquire) > 0 && now_s() - t0 < ELAS_TRI_SPIN_S) _mm_pause();
    std::unique_lock<std::mutex> g(t->batch.m);
    t->batch.cv.wait(g, [&] { return t->batch.remaining == 0; });
}

void elas_tri_release(ElasTriTicket* t) {
    if (!t) return;
    elas_tri_wait(t);
    delete t;
}
