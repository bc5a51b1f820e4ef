// elas_triangle.h -- host Delaunay jobs over independent Triangle instances.
#pragma once
#include <stdint.h>
#include <vector>

#define ELAS_TRI_INSTANCES 16

// Runs task(arg, i) for every i in [0, ntasks) and returns when all are done
// (the tasks may run on other threads).
typedef void (*ElasParallelFor)(void* ctx, int ntasks, void (*task)(void* arg, int i), void* arg);

struct ElasTriJob {
    const float* pts;            // [n][2] support points
    int n;
    std::vector<int32_t>* out;   // [ntri][3] zero-based triangles
    int result;                  // ntri, or -1 for fewer than 3 points
    // filled by the worker: seconds (steady clock) around the Triangle call
    double t_begin = 0, t_end = 0;
};

// The vertex order Triangle's divide-and-conquer builds for a point set
// (input indices), when the coordinates are bounded integers and no two points
// coincide; false (arr unused) otherwise.
bool elas_tri_arrangement(const float* p, int n, std::vector<int32_t>& arr);

// Runs every job (each a "zQB" Triangle call) on up to max_workers of the
// process-wide Triangle workers; returns when all are done.
void elas_triangulate_jobs(ElasTriJob* jobs, int n, int max_workers);

// Largest split of one mesh's divide-and-conquer merge over the pool's
// threads: up to 2^max_depth concurrent subtrees (0: serial Triangle; -1:
// automatic, subtrees of at least a few hundred points, at most 8).  The
// meshes are the same under every setting.  elas_tri_split_depth() is the
// depth a mesh of n points gets now.
void elas_tri_set_split(int max_depth);
int elas_tri_split_depth(int n);

// Non-blocking form: queues the jobs and returns a ticket; elas_tri_done()
// tells whether they have all finished, elas_tri_wait() blocks until then.
// The jobs must stay alive until done; elas_tri_release() frees a finished
// ticket.
typedef struct ElasTriTicket ElasTriTicket;
ElasTriTicket* elas_tri_submit(ElasTriJob* jobs, int n);
bool elas_tri_done(ElasTriTicket* t);
void elas_tri_wait(ElasTriTicket* t);
void elas_tri_release(ElasTriTicket* t);
