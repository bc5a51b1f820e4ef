// elas_triangle.h -- host Delaunay jobs over independent Triangle instances.
#pragma once
#include <stdint.h>
#include <vector>

#define ELAS_TRI_INSTANCES 16

#define ELAS_RASTER_TILE 32

struct ElasTriJob {
    const float* pts;            // [n][2] support points
    int n;
    std::vector<int32_t>* out;   // [ntri][3] zero-based triangles
    int result;                  // ntri, or -1 for fewer than 3 points
    // optional tile binning of the mesh (done by the same worker): for every
    // ELAS_RASTER_TILE^2 pixel tile of a W x H image, the ascending list of
    // triangles whose pixel bounding box meets it
    int W, H;
    std::vector<int32_t>* tile_off;   // [ntiles + 1]
    std::vector<int32_t>* tile_list;  // [tile_off[ntiles]]
    // filled by the worker: seconds (steady clock) around Triangle and binning
    double t_begin = 0, t_tri = 0, t_end = 0;
};

// Tile binning of triangles with corners pts (pixel coordinates).
void elas_bin_triangles(const float* pts, const int32_t* tri, int ntri, int W, int H,
                        std::vector<int32_t>& off, std::vector<int32_t>& list);

// The (x, y) order of a point set when it can be handed to Triangle in place
// of the caller's order with the same mesh: integer x, no duplicates.  order
// maps sorted position to input index.  False (order unused) otherwise.
bool elas_tri_sorted_order(const float* p, int n, std::vector<int32_t>& order, std::vector<float>& sorted);

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
