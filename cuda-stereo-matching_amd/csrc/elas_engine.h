// elas_engine.h -- host engine behind the C ABI: one context = one device,
// one pair of HIP streams, one HBM arena, pinned staging, stage timers.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "elas_common.h"
#include "elas_hip.h"
#include "elas_triangle.h"

// One pair of a batch.  Buffers are host pointers unless dev_* is set.
struct ElasPairSpec {
    int32_t W, H, dmax;
    const uint8_t* left;
    const uint8_t* right;
    float* D1;
    float* D2;
    bool dev_in, dev_out;
};

struct elas_ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;   // main pipeline
    hipStream_t stream2 = nullptr;  // disparity grid, overlapped with host Triangle
    elas_params prm;
    ElasKParams kp;
    bool debug = false;
    int tri_workers = ELAS_TRI_INSTANCES;
    std::string err;

    // device memory
    void* arena = nullptr;
    size_t arena_bytes = 0;
    void* table = nullptr;  // ElasPair[n] + prefix[S_NUM_STAGES][n+1] + nsp[n]
    size_t table_bytes = 0;
    // pinned host staging
    void* pin = nullptr;
    size_t pin_bytes = 0;
    // raster tile offsets + lists (sizes known only after Triangle)
    void* blist = nullptr;
    size_t blist_bytes = 0;
    void* blist_pin = nullptr;
    size_t blist_pin_bytes = 0;

    std::vector<ElasPair> pairs;                    // host copy of the device table
    std::vector<int32_t> prefix;                    // [S_NUM_STAGES][n+1]
    std::vector<ElasPairSpec> specs;                // last call
    std::vector<int32_t> nsp, ntri0, ntri1, status; // last call, per pair
    hipEvent_t ev[16] = {};
    hipEvent_t ev_grid = nullptr;
    float times[ELAS_T_COUNT] = {};

    int run(const ElasPairSpec* specs, int n, int32_t* statuses);
    int fetch(int stage, void* dst, size_t bytes, size_t* stage_bytes);
    int fail(int code, const std::string& msg) { err = msg; return code; }
};

// Parameter checks and derived kernel parameters (prior table, key bias).
int elas_make_kparams(const elas_params* p, ElasKParams* kp, std::string* err);
