// elas_triangle.h -- host Delaunay jobs over independent Triangle instances.
#pragma once
#include <stdint.h>
#include <vector>

#define ELAS_TRI_INSTANCES 16

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

// Non-blocking form: queues the jobs and returns a ticket; elas_tri_done()
// tells whether they have all finished, elas_tri_wait() blocks until then.
// The jobs must stay alive until done; elas_tri_release() frees a finished
// ticket.
typedef struct ElasTriTicket ElasTriTicket;
ElasTriTicket* elas_tri_submit(ElasTriJob* jobs, int n);
bool elas_tri_done(ElasTriTicket* t);
void elas_tri_wait(ElasTriTicket* t);
void elas_tri_release(ElasTriTicket* t);
