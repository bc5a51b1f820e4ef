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
};

// Runs every job (each a "zQB" Triangle call) on up to max_workers of the
// process-wide Triangle workers; returns when all are done.
void elas_triangulate_jobs(ElasTriJob* jobs, int n, int max_workers);
