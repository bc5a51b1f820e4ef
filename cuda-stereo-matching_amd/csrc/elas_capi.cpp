// elas_capi.cpp -- extern "C" entry points declared in include/elas_hip.h.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "elas_engine.h"
#include "elas_hip.h"
#include "elas_launch.h"
#include "elas_triangle.h"

extern "C" {

// Effective values of the reference kernels (src/Kernels/KernelsCommon.h:8-27);
// post-processing off because the reference pipeline ends at the LR check.
void elas_params_default(elas_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->disp_min = 0;
    p->disp_max = 255;
    p->support_threshold = 0.85f;
    p->support_texture = 10;
    p->candidate_stepsize = 5;
    p->incon_window_size = 5;
    p->incon_threshold = 5;
    p->incon_min_support = 5;
    p->add_corners = false;
    p->grid_size = 20;
    p->beta = 0.02f;
    p->gamma = 3.0f;
    p->sigma = 1.0f;
    p->sradius = 2.0f;
    p->match_texture = 0;
    p->lr_threshold = 2;
    p->speckle_sim_threshold = 1.0f;
    p->speckle_size = 0;
    p->ipol_gap_width = 0;
    p->filter_median = false;
    p->filter_adaptive_mean = false;
    p->postprocess_only_left = false;
    p->subsampling = false;
}

// The initialisers written in Elas::parameters() (src/Elas.h:42-66), which
// the reference kernels never read.
void elas_params_declared(elas_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->disp_min = 0;
    p->disp_max = 255;
    p->support_threshold = 0.95f;
    p->support_texture = 10;
    p->candidate_stepsize = 5;
    p->incon_window_size = 5;
    p->incon_threshold = 5;
    p->incon_min_support = 5;
    p->add_corners = true;
    p->grid_size = 20;
    p->beta = 0.02f;
    p->gamma = 5.0f;
    p->sigma = 1.0f;
    p->sradius = 3.0f;
    p->match_texture = 0;
    p->lr_threshold = 2;
    p->speckle_sim_threshold = 1.0f;
    p->speckle_size = 200;
    p->ipol_gap_width = 5000;
    p->filter_median = true;
    p->filter_adaptive_mean = false;
    p->postprocess_only_left = false;
    p->subsampling = false;
}

int elas_hip_create(const elas_params* params, int device, elas_ctx** out) {
    if (!out) return ELAS_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    {
        // parameters first: a request the library cannot honour is refused
        // the same way with or without a device
        elas_params p;
        if (params) p = *params;
        else elas_params_default(&p);
        ElasKParams kp;
        std::string e;
        const int st = elas_make_kparams(&p, &kp, &e);
        if (st != ELAS_OK) return st;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return ELAS_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev) return ELAS_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return ELAS_ERR_NO_DEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return ELAS_ERR_NO_DEVICE;
    elas_ctx* c = new (std::nothrow) elas_ctx();
    if (!c) return ELAS_ERR_OUT_OF_MEMORY;
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    c->max_lds = std::min<size_t>(prop.sharedMemPerBlock, 163840);
    elas_params p;
    if (params) p = *params;
    else elas_params_default(&p);
    int st = elas_make_kparams(&p, &c->kp, &c->err);
    if (st != ELAS_OK) {
        delete c;
        return st;
    }
    c->prm = p;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return ELAS_ERR_HIP;
    }
    if (elas_dense_allow_lds(c->max_lds) != 0) c->max_lds = 65536;
    st = elas_hip_set_pipeline(c, 3, 0);
    if (st == ELAS_OK && c->fixed_init() != ELAS_OK) {
        // every other entry point works without it; the no-allocation one
        // reports the reason (kept in fixed_err)
        c->fixed_err = c->err;
        c->err.clear();
        if (c->fixed) {
            c->slot_free(*c->fixed);
            delete c->fixed;
            c->fixed = nullptr;
        }
        (void)hipGetLastError();
    }
    if (st != ELAS_OK) {
        elas_hip_destroy(c);
        return st;
    }
    *out = c;
    return ELAS_OK;
}

#define ELAS_NOT_SERVING(c)                                                           \
    do {                                                                              \
        if ((c)->server) return (c)->fail(ELAS_ERR_BUSY, "context is serving requests"); \
    } while (0)

int elas_hip_destroy(elas_ctx* c) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    elas_srv_stop(c);
    for (ElasSlot* S : c->slots) {
        c->slot_free(*S);
        delete S;
    }
    c->slots.clear();
    if (c->fixed) {
        c->slot_free(*c->fixed);
        delete c->fixed;
    }
    delete c;
    return ELAS_OK;
}

int elas_hip_set_graphs(elas_ctx* c, int32_t on) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    c->use_graphs = on != 0;
    return ELAS_OK;
}

int elas_hip_set_pipeline(elas_ctx* c, int32_t slots, int32_t chunk) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    if (slots < 1 || slots > 8 || chunk < 0) {
        c->err = "need 1 <= slots <= 8 and chunk >= 0";
        return ELAS_ERR_INVALID_ARGUMENT;
    }
    while ((int)c->slots.size() > slots) {
        ElasSlot* S = c->slots.back();
        if (c->last == S) c->last = nullptr;
        c->slot_free(*S);
        delete S;
        c->slots.pop_back();
    }
    while ((int)c->slots.size() < slots) {
        ElasSlot* S = new (std::nothrow) ElasSlot();
        if (!S) return ELAS_ERR_OUT_OF_MEMORY;
        c->slots.push_back(S);
        const int st = c->slot_init(*S);
        if (st != ELAS_OK) return st;
    }
    c->chunk = chunk;
    return ELAS_OK;
}

int elas_hip_set_params(elas_ctx* c, const elas_params* p) {
    if (!c || !p) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    ElasKParams kp;
    int st = elas_make_kparams(p, &kp, &c->err);
    if (st != ELAS_OK) return st;
    c->prm = *p;
    c->kp = kp;
    c->param_gen++;
    c->tuned.clear();  // tilings depend on grid_size
    return ELAS_OK;
}

const char* elas_hip_last_error(const elas_ctx* c) {
    if (!c) return "null context";
    static thread_local std::string copy;
    std::lock_guard<std::mutex> g(const_cast<elas_ctx*>(c)->err_m);
    copy = c->err;
    return copy.c_str();
}

void* elas_hip_stream(elas_ctx* c) { return c && !c->slots.empty() ? (void*)c->slots[0]->stream : nullptr; }

int elas_hip_process(elas_ctx* c, const uint8_t* left, const uint8_t* right, float* D1, float* D2,
                     const int32_t* dims) {
    if (!c || !dims) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    ElasPairSpec s{dims[0], dims[1], c->prm.disp_max, left, right, D1, D2, false, false, -1};
    return c->run(&s, 1, nullptr);
}

int elas_hip_process_device(elas_ctx* c, const uint8_t* d_left, const uint8_t* d_right, float* d_D1,
                            float* d_D2, int32_t width, int32_t height) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    ElasPairSpec s{width, height, c->prm.disp_max, d_left, d_right, d_D1, d_D2, true, true, -1};
    return c->run(&s, 1, nullptr);
}

int elas_hip_process_device_noalloc(elas_ctx* c, const uint8_t* d_left, const uint8_t* d_right, float* d_D1,
                                    float* d_D2, int32_t width, int32_t height) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    ElasPairSpec s{width, height, c->prm.disp_max, d_left, d_right, d_D1, d_D2, true, true, -1};
    return c->run_fixed(s);
}

static int batch(elas_ctx* c, int32_t n, const uint8_t* const* left, const uint8_t* const* right,
                 float* const* D1, float* const* D2, const int32_t* dims, int32_t* statuses, bool dev) {
    if (!c || n <= 0 || !left || !right || !D1 || !D2 || !dims) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    std::vector<ElasPairSpec> v(n);
    for (int i = 0; i < n; i++) {
        const int dm = dims[3 * i + 2] < 0 ? c->prm.disp_max : dims[3 * i + 2];
        v[i] = ElasPairSpec{dims[3 * i], dims[3 * i + 1], dm, left[i], right[i], D1[i], D2[i], dev, dev, -1};
    }
    return c->run_batch(v.data(), n, statuses, c->chunk);
}

int elas_hip_process_batch(elas_ctx* c, int32_t n, const uint8_t* const* left, const uint8_t* const* right,
                           float* const* D1, float* const* D2, const int32_t* dims, int32_t* statuses) {
    return batch(c, n, left, right, D1, D2, dims, statuses, false);
}

int elas_hip_process_batch_device(elas_ctx* c, int32_t n, const uint8_t* const* d_left,
                                  const uint8_t* const* d_right, float* const* d_D1, float* const* d_D2,
                                  const int32_t* dims, int32_t* statuses) {
    return batch(c, n, d_left, d_right, d_D1, d_D2, dims, statuses, true);
}

int elas_hip_set_debug(elas_ctx* c, int keep) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    c->debug = (keep & 1) != 0;
    c->tiny_lists = (keep & 2) != 0;
    c->gate_b = (keep & 4) == 0;
    c->gate_ticks = (keep & 8) ? 1ull : 3000000000ull;
    c->param_gen++;
    return ELAS_OK;
}

int elas_hip_debug_fetch(elas_ctx* c, int stage, void* dst, size_t bytes, size_t* stage_bytes) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    return c->fetch(stage, dst, bytes, stage_bytes);
}

int elas_hip_shape_info(const elas_ctx* c, elas_shape_info* info) {
    if (!c || !info) return ELAS_ERR_INVALID_ARGUMENT;
    memset(info, 0, sizeof(*info));
    if (c->server || !c->last || c->last->pairs.empty()) return ELAS_ERR_INVALID_ARGUMENT;
    const ElasSlot& S = *c->last;
    const ElasPair& P = S.pairs[0];
    info->width = P.W;
    info->height = P.H;
    info->can_width = P.cw;
    info->can_height = P.ch;
    info->grid_width = P.gw;
    info->grid_height = P.gh;
    info->num_support = S.nsp[0];
    info->num_tri0 = S.ntri0[0];
    info->num_tri1 = S.ntri1[0];
    info->status = S.status[0];
    return ELAS_OK;
}

int elas_hip_stage_times(const elas_ctx* c, float* ms, int32_t n) {
    if (!c || !ms || c->server) return ELAS_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < n && i < ELAS_T_COUNT; i++) ms[i] = c->last ? c->last->times[i] : 0.0f;
    return ELAS_OK;
}

int elas_hip_set_autotune(elas_ctx* c, int32_t on) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    c->autotune = on != 0;
    return ELAS_OK;
}

int elas_hip_set_dense_tiles(elas_ctx* c, int32_t cells_per_wave, int32_t steps_per_chunk) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    if (cells_per_wave < 0 || cells_per_wave > 2 || steps_per_chunk < 0 ||
        (cells_per_wave == 0 && steps_per_chunk != 0)) {
        c->err = "need cells_per_wave in {0, 1, 2}, steps_per_chunk >= 0, and both 0 for automatic tiling";
        return ELAS_ERR_INVALID_ARGUMENT;
    }
    c->force_cpw = cells_per_wave;
    c->force_spc = steps_per_chunk;
    return ELAS_OK;
}

int elas_hip_dense_tiles(elas_ctx* c, int32_t W, int32_t H, int32_t disp_max, int32_t* cells_per_wave,
                         int32_t* steps_per_chunk, float* tuned_us) {
    if (!c || W < 1 || H < 1 || disp_max < 0) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    const ElasTiles t = c->tiles(elas_dims(c->prm, W, H, disp_max), W, H, disp_max);
    if (cells_per_wave) *cells_per_wave = t.cpw;
    if (steps_per_chunk) *steps_per_chunk = t.spc;
    if (tuned_us) *tuned_us = t.us;
    return ELAS_OK;
}

int elas_hip_set_desc_rows(elas_ctx* c, int32_t rows) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    if (rows != 0 && rows != 6 && rows != 12 && rows != 18) {
        c->err = "descriptor strip rows must be 6, 12, 18 or 0 (automatic)";
        return ELAS_ERR_INVALID_ARGUMENT;
    }
    c->force_dsh = rows;
    return ELAS_OK;
}

int elas_hip_desc_rows(elas_ctx* c, int32_t W, int32_t H, int32_t disp_max, int32_t* rows, float* tuned_us) {
    if (!c || W < 1 || H < 1 || disp_max < 0) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    const ElasTiles t = c->tiles(elas_dims(c->prm, W, H, disp_max), W, H, disp_max);
    if (rows) *rows = t.dsh;
    if (tuned_us) *tuned_us = t.desc_us;
    return ELAS_OK;
}

int elas_hip_time_kernel(elas_ctx* c, int32_t kernel, int32_t reps, float* avg_ms) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    ELAS_NOT_SERVING(c);
    return c->time_kernel(kernel, reps, avg_ms);
}

int elas_hip_serve_start(elas_ctx* c, const elas_serve_config* cfg) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    elas_serve_config d{0, 0, 0};
    if (cfg) d = *cfg;
    return elas_srv_start(c, d.hbm_budget_bytes, d.max_chunk_pairs, d.slots);
}

int elas_hip_submit(elas_ctx* c, const uint8_t* left, const uint8_t* right, float* D1, float* D2, int32_t width,
                    int32_t height, int32_t disp_max, int32_t flags, int64_t* ticket) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    const bool dev = (flags & 1) != 0;
    ElasPairSpec s{width, height, disp_max < 0 ? c->prm.disp_max : disp_max, left, right, D1, D2, dev, dev, -1};
    return elas_srv_submit(c, s, ticket);
}

int elas_hip_poll(elas_ctx* c, int64_t ticket, int32_t* status, float* latency_ms) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    return elas_srv_poll(c, ticket, status, latency_ms);
}

int elas_hip_serve_stop(elas_ctx* c) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    return elas_srv_stop(c);
}

int elas_hip_triangulate(const float* pts, int32_t n, int32_t* tri, int64_t cap, int64_t* ntri) {
    if ((!pts && n > 0) || n < 0 || cap < 0 || (cap > 0 && !tri) || !ntri) return ELAS_ERR_INVALID_ARGUMENT;
    std::vector<int32_t> out;
    ElasTriJob job{};
    job.pts = pts;
    job.n = n;
    job.out = &out;
    elas_triangulate_jobs(&job, 1, ELAS_TRI_INSTANCES);  // on this thread, the merge split over the pool
    if (job.result < 0) {
        *ntri = 0;
        return ELAS_ERR_TOO_FEW_POINTS;  // n < 3, or every point the same point
    }
    *ntri = job.result;
    memcpy(tri, out.data(), sizeof(int32_t) * 3 * (size_t)std::min<int64_t>(cap, job.result));
    return ELAS_OK;
}

int elas_hip_set_tri_split(int32_t max_depth) {
    if (max_depth < -1 || max_depth > 12) return ELAS_ERR_INVALID_ARGUMENT;
    elas_tri_set_split(max_depth);
    return ELAS_OK;
}

int elas_hip_tri_split_depth(int32_t n) { return elas_tri_split_depth(n); }

int elas_hip_serve_stats(elas_ctx* c, elas_serve_stats* st) {
    if (!c) return ELAS_ERR_INVALID_ARGUMENT;
    return elas_srv_stats(c, st);
}

}  // extern "C"
