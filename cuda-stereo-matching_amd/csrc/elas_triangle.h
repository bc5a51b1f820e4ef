// elas_triangle.h -- host Delaunay jobs over independent Triangle instances.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <vector>

#define ELAS_TRI_INSTANCES 16

// Runs task(arg, i) for every i in [0, ntasks) and returns when all are done
// (the tasks may run on other threads).
typedef void (*ElasParallelFor)(void* ctx, int ntasks, void (*task)(void* arg, int i), void* arg);

struct ElasTriJob {
    const float* pts;            // [n][2] support points
    int n;
    std::vector<int32_t>* out;   // [ntri][3] zero-based triangles, unless dst is set
    int32_t* dst = nullptr;      // else: written here directly, if ntri <= cap
    int64_t cap = 0;             // capacity of dst in triangles
    bool dst16 = false;          // dst takes 16-bit corner indices ([ntri][3] uint16; n <= 65536)
    int result;                  // ntri, or -1 for fewer than 3 distinct points
    // filled by the worker: seconds (steady clock) around the Triangle call
    double t_begin = 0, t_end = 0;
};

// Triangle's divide-and-conquer vertex order (divconqdelaunay(),
// src/triangle.cpp:6164-6230): the points sorted by (x, y), halved, and each
// half cut alternately along y and x down to 2- and 3-point leaves.  ly holds
// the points in (y, x) order (and lx, for pack(), in (x, y) order); cut() works on a node
// [lo, lo + size): a cut along `axis` keeps the size / 2 smallest points of
// that order on the left (as vertexmedian() does, src/triangle.cpp:
// 5517-5574) and splits the other list stably, so both children keep both
// orders; 2- and 3-point nodes are leaves in (x, y) order (alternateaxes(),
// :5586-5605) and go to arr.  `levels` bounds the depth of the cuts made now
// (-1: down to the leaves): the top of the tree can be cut on one thread and
// each subtree below it on another, since subtrees touch disjoint ranges.
// The cuts work on px and py, the two orders as entries (x rank << 32 |
// y rank), where a point's x rank is its place in (x, y) order and its y
// rank its place in ly: a point belongs to the left of a cut along x exactly
// when its x rank is at most that of the node's `divider`-th point in x
// order, so a cut reads its entries in sequence and looks nothing up; a leaf
// maps its entries to points through ly.  pack() builds px and py from lx
// and ly.
struct ElasTriOrder {
    int lohi[4];  // integer orders: min x, max x, min y, max y
    std::vector<int32_t> lx, ly, arr, rank;
    // entries: 64-bit (32-bit ranks), or with `narrow` (at most 65536
    // points) 32-bit (16-bit ranks) in qx, qy, ql, qr, which halves the
    // memory the cuts stream through
    bool narrow = false;
    std::vector<uint64_t> px, py, tl, tr;
    std::vector<uint32_t> qx, qy, ql, qr;
    void resize(int n, bool narrow_entries);
    void pack();
    void cut(int lo, int size, int axis, int levels);
};

// The order (lx, ly, px, py) of a point set by counting sorts, when the
// coordinates are bounded integers and no two points coincide (support
// points); false otherwise (which duplicate Triangle keeps then depends on
// its own sort).
bool elas_tri_order_integer(const float* p, int n, ElasTriOrder& o);
// lx and ly of `count` distinct points given in (x, y) order (positions).
void elas_tri_order_sorted(const float* xy, int count, ElasTriOrder& o);

// Triangle's allocator in this library (its trimalloc() and trifree() call
// these): blocks of 64 KiB and more are kept by power-of-two size class and
// handed out again, first from the freeing thread's own list (no lock, lines
// likely still in its cache), then from a process-wide cache, so the pools
// of one call reuse memory that earlier calls already touched instead of
// faulting in fresh pages from several threads at once; smaller blocks go
// to malloc.
void* elas_tri_block_alloc(size_t bytes);
void elas_tri_block_free(void* p);

// Runs every job (each a "zQB" Triangle call) on up to max_workers of the
// process-wide Triangle workers; returns when all are done.
void elas_triangulate_jobs(ElasTriJob* jobs, int n, int max_workers);

// Largest split of one mesh's divide-and-conquer merge over the pool's
// threads: up to 2^max_depth concurrent subtrees (0: serial Triangle; -1:
// automatic, subtrees of at least 128 points, at most 8).  The
// meshes are the same under every setting.  elas_tri_split_depth() is the
// depth a mesh of n points gets now.
void elas_tri_set_split(int max_depth);
int elas_tri_split_depth(int n);

// A single-pair call about to submit its two meshes (of about n0 and n1
// points): the workers that will take them warm their order scratch
// meanwhile.  Optional; the meshes are the same either way.
void elas_tri_prewarm(int n0, int n1);

// Non-blocking form: queues the jobs and returns a ticket; elas_tri_done()
// tells whether they have all finished, elas_tri_wait() blocks until then.
// The jobs must stay alive until done; elas_tri_release() frees a finished
// ticket.
typedef struct ElasTriTicket ElasTriTicket;
ElasTriTicket* elas_tri_submit(ElasTriJob* jobs, int n);
bool elas_tri_done(ElasTriTicket* t);
void elas_tri_wait(ElasTriTicket* t);
void elas_tri_release(ElasTriTicket* t);
