// elas_engine.cpp -- host orchestration of the ELAS path on one gfx950 GPU.
//
// Stage order, buffer initial values and the host Delaunay excursion follow
// Elas::process (src/Elas.cu:19-120).  Differences of structure, none of
// arithmetic:
//   * every stage is one launch over a whole batch of pairs (a single process
//     call is a batch of one);
//   * device buffers live in one arena that grows on demand and is reused
//     across calls (the reference allocates ~20 buffers per call and leaks
//     its scan buffers, Elas.cu:349-371);
//   * the disparity grid (Elas.cu:65) runs on a second stream while the host
//     triangulates, and the left and right meshes are triangulated in
//     parallel by independent Triangle instances;
//   * buffers the reference initialises with thrust::fill (-1 for D1/D2 and
//     the plane map, 0 for the grid, Elas.cu:94-95,415-420,488) are written in
//     full by the kernels themselves, or cleared by fill kernels.
#include "elas_engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "elas_launch.h"
#include "elas_triangle.h"

#define HIPCHECK(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return fail(ELAS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

size_t al(size_t x) { return (x + 255) & ~size_t(255); }

struct Dims {
    int32_t cw, ch, gw, gh, words, nblk, tricap;
    size_t npx, ncan, ncell;
};

Dims dims_of(const elas_params& p, int W, int H, int dmax) {
    Dims d;
    d.cw = W / p.candidate_stepsize;
    d.ch = H / p.candidate_stepsize;
    d.gw = (int32_t)ceil((float)W / (float)p.grid_size);  // Elas.cu:411-412
    d.gh = (int32_t)ceil((float)H / (float)p.grid_size);
    d.words = (dmax + 32) / 32;
    d.npx = (size_t)W * H;
    d.ncan = (size_t)d.cw * d.ch;
    d.ncell = (size_t)d.gw * d.gh;
    d.nblk = (int32_t)((d.ncan + 255) / 256);
    d.tricap = (int32_t)(2 * d.ncan + 8);
    return d;
}

int dense_rows(int gs) { return std::max(1, std::min(64 / gs, gs)); }

// LDS of a dense tile: R rows x (16*cpw cells x gs + dmax) columns x 16 B
size_t dense_lds(int gs, int dmax, int cpw) {
    return (size_t)dense_rows(gs) * (size_t)(ELAS_DENSE_WAVES * cpw * gs + dmax) * 16;
}

int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace

int elas_make_kparams(const elas_params* p, ElasKParams* kp, std::string* err) {
    auto bad = [&](const char* m) {
        *err = m;
        return (int)ELAS_ERR_INVALID_ARGUMENT;
    };
    if (p->candidate_stepsize < 1) return bad("candidate_stepsize must be >= 1");
    if (p->grid_size < 2 || p->grid_size > 64) return bad("grid_size must be in [2, 64]");
    if (p->disp_min < 0 || p->disp_max < p->disp_min || p->disp_max > ELAS_MAX_DISP)
        return bad("need 0 <= disp_min <= disp_max <= 511");
    if (p->incon_window_size < 0 || p->incon_window_size > 64) return bad("incon_window_size must be in [0, 64]");
    if (!(p->beta > 0.0f) || !(p->gamma > 0.0f) || !(p->sigma > 0.0f) || !(p->sradius >= 0.0f))
        return bad("beta, gamma, sigma must be > 0 and sradius >= 0");
    if (!(p->support_threshold == p->support_threshold)) return bad("support_threshold is NaN");
    memset(kp, 0, sizeof(*kp));
    kp->support_threshold = p->support_threshold;
    kp->support_texture = p->support_texture;
    kp->incon_window = p->incon_window_size;
    kp->incon_threshold = p->incon_threshold;
    kp->incon_min_support = p->incon_min_support;
    kp->lr_threshold = p->lr_threshold;
    kp->match_texture = p->match_texture;
    // plane radius, KerComputeDisparity.cu:21,194
    const int pr = (int)std::max((float)ceilf(p->sigma * p->sradius), 2.0f);
    if (pr >= ELAS_MAX_P) return bad("ceil(sigma * sradius) must be < 16");
    kp->plane_radius = pr;
    // prior P(delta), KerComputeDisparity.cu:249-251,269-271
    int pmin = 0;
    bool allneg = true;
    for (int k = 0; k <= pr; k++) {
        const float two_sigma_squared = 2 * p->sigma * p->sigma;
        const float e = expf((float)(-k * k) / two_sigma_squared);
        const float num = -logf(p->gamma + e) + logf(p->gamma);
        const int P = (int)((double)num / (double)p->beta);
        kp->P[k] = P;
        pmin = std::min(pmin, P);
        allneg = allneg && P < 0;
    }
    if (pmin < -(1 << 18)) return bad("prior too strong (beta too small)");
    kp->key_bias = -pmin;
    kp->all_p_negative = allneg ? 1 : 0;
    return ELAS_OK;
}

int elas_ctx::run(const ElasPairSpec* sp, int n, int32_t* statuses) {
    using clk = std::chrono::steady_clock;
    const auto t_start = clk::now();
    if (n <= 0 || !sp) return fail(ELAS_ERR_INVALID_ARGUMENT, "empty batch");
    if (prm.speckle_size > 0 || prm.ipol_gap_width > 0 || prm.filter_median || prm.filter_adaptive_mean)
        return fail(ELAS_ERR_UNSUPPORTED, "post-processing filters are not available on the GPU path yet");
    HIPCHECK(hipSetDevice(device));
    specs.assign(sp, sp + n);
    std::vector<Dims> dd(n);
    size_t arena_need = 0, pin_sp = 0, pin_tri = 0;
    std::vector<size_t> base(n);
    for (int i = 0; i < n; i++) {
        const ElasPairSpec& s = sp[i];
        if (s.W < 1 || s.H < 1 || s.W > 16384 || s.H > 16384 || (size_t)s.W * s.H > (size_t)1 << 27)
            return fail(ELAS_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": dims out of range");
        if (!s.left || !s.right || !s.D1 || !s.D2)
            return fail(ELAS_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": null buffer");
        if (s.dmax < std::max(prm.disp_min, 0) || s.dmax > ELAS_MAX_DISP)
            return fail(ELAS_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": disp_max out of range");
        if (dense_lds(prm.grid_size, s.dmax, 1) > 65536)
            return fail(ELAS_ERR_UNSUPPORTED, "grid_size/disp_max exceed the dense LDS tile");
        if (2 * 16 * (size_t)(63 * prm.candidate_stepsize + 2 * s.dmax + 5) > 65536)
            return fail(ELAS_ERR_UNSUPPORTED, "candidate_stepsize/disp_max exceed the support LDS rows");
        Dims d = dims_of(prm, s.W, s.H, s.dmax);
        dd[i] = d;
        size_t b = 0;
        if (!s.dev_in) b += 2 * al(d.npx);
        b += 2 * al(16 * d.npx);
        b += 4 * al(4 * d.ncan) + al(4 * (size_t)d.nblk) + 2 * al(8 * d.ncan);
        b += 4 * al(4 * d.ncell * d.words);
        if (debug) b += 2 * al(4 * d.ncell * (size_t)(s.dmax + 2)) + 2 * al(4 * d.npx);
        b += 2 * al(12 * (size_t)d.tricap) + 2 * al(4 * ELAS_TRIREC * (size_t)d.tricap);
        b += 2 * al(4 * d.npx) + 2 * al(4 * d.npx);  // prior, draw
        if (!s.dev_out) b += 2 * al(4 * d.npx);
        base[i] = arena_need;
        arena_need += b;
        pin_sp += 2 * 8 * d.ncan;
        pin_tri += 2 * 12 * (size_t)d.tricap;
    }
    int maxd = 0;
    for (int i = 0; i < n; i++) maxd = std::max(maxd, sp[i].dmax);
    const int cpw = dense_lds(prm.grid_size, maxd, 2) <= 65536 ? 2 : 1;
    ElasKParams kpl = kp;
    kpl.dense_cpw = cpw;
    if (arena_need > arena_bytes) {
        if (arena) (void)hipFree(arena);
        arena = nullptr;
        arena_bytes = 0;
        const size_t want = arena_need + arena_need / 4;
        if (hipMalloc(&arena, want) != hipSuccess) {
            arena = nullptr;
            (void)hipGetLastError();
            return fail(ELAS_ERR_OUT_OF_MEMORY, "arena allocation of " + std::to_string(want) + " bytes failed");
        }
        arena_bytes = want;
    }
    // two device tables (phase A / phase B): pairs, prefixes, nsp
    const size_t tab_pairs = al(sizeof(ElasPair) * n);
    const size_t tab_prefix = al(sizeof(int32_t) * S_NUM_STAGES * (n + 1));
    const size_t tab_one = tab_pairs + tab_prefix + al(4 * (size_t)n);
    if (2 * tab_one > table_bytes) {
        if (table) (void)hipFree(table);
        table = nullptr;
        table_bytes = 0;
        if (hipMalloc(&table, 2 * tab_one) != hipSuccess) {
            table = nullptr;
            (void)hipGetLastError();
            return fail(ELAS_ERR_OUT_OF_MEMORY, "table allocation failed");
        }
        table_bytes = 2 * tab_one;
    }
    const size_t pin_need = 2 * tab_one + al(pin_sp) + al(pin_tri);
    if (pin_need > pin_bytes) {
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        pin_bytes = 0;
        if (hipHostMalloc(&pin, pin_need + pin_need / 4, hipHostMallocDefault) != hipSuccess) {
            pin = nullptr;
            (void)hipGetLastError();
            return fail(ELAS_ERR_OUT_OF_MEMORY, "pinned staging allocation failed");
        }
        pin_bytes = pin_need + pin_need / 4;
    }
    char* A = (char*)arena;
    char* T = (char*)table;
    char* PIN = (char*)pin;
    // device table layout
    ElasPair* dpairs[2] = {(ElasPair*)T, (ElasPair*)(T + tab_one)};
    int32_t* dprefix[2] = {(int32_t*)(T + tab_pairs), (int32_t*)(T + tab_one + tab_pairs)};
    int32_t* dnsp = (int32_t*)(T + tab_pairs + tab_prefix);
    ElasPair* hpairs = (ElasPair*)PIN;
    int32_t* hprefix = (int32_t*)(PIN + tab_pairs);
    int32_t* hnsp = (int32_t*)(PIN + tab_pairs + tab_prefix);
    ElasPair* hpairs_b = (ElasPair*)(PIN + tab_one);
    int32_t* hprefix_b = (int32_t*)(PIN + tab_one + tab_pairs);
    float* hsp = (float*)(PIN + 2 * tab_one);
    int32_t* htri = (int32_t*)(PIN + 2 * tab_one + al(pin_sp));

    pairs.assign(n, ElasPair());
    prefix.assign((size_t)S_NUM_STAGES * (n + 1), 0);
    nsp.assign(n, 0);
    ntri0.assign(n, 0);
    ntri1.assign(n, 0);
    status.assign(n, ELAS_OK);
    std::vector<size_t> sp_off(n), tri_off(n);
    size_t so = 0, to = 0;
    for (int i = 0; i < n; i++) {
        const ElasPairSpec& s = sp[i];
        const Dims& d = dd[i];
        ElasPair& P = pairs[i];
        memset(&P, 0, sizeof(P));
        P.W = s.W; P.H = s.H; P.dmin = prm.disp_min; P.dmax = s.dmax;
        P.cw = d.cw; P.ch = d.ch; P.gw = d.gw; P.gh = d.gh; P.words = d.words;
        P.step = prm.candidate_stepsize; P.gsize = prm.grid_size; P.nblk = d.nblk;
        char* c = A + base[i];
        auto take = [&](size_t bytes) { char* r = c; c += al(bytes); return r; };
        if (!s.dev_in) { P.img[0] = (uint8_t*)take(d.npx); P.img[1] = (uint8_t*)take(d.npx); }
        else { P.img[0] = s.left; P.img[1] = s.right; }
        for (int k = 0; k < 2; k++) P.desc[k] = (uint8_t*)take(16 * d.npx);
        for (int k = 0; k < 4; k++) P.can[k] = (int32_t*)take(4 * d.ncan);
        P.blkcnt = (int32_t*)take(4 * (size_t)d.nblk);
        for (int k = 0; k < 2; k++) P.sp[k] = (float*)take(8 * d.ncan);
        P.nsp = dnsp + i;
        for (int k = 0; k < 2; k++) P.gflag[k] = (uint32_t*)take(4 * d.ncell * d.words);
        for (int k = 0; k < 2; k++) P.gmask[k] = (uint32_t*)take(4 * d.ncell * d.words);
        if (debug) {
            for (int k = 0; k < 2; k++) P.glist[k] = (int32_t*)take(4 * d.ncell * (size_t)(s.dmax + 2));
            for (int k = 0; k < 2; k++) P.mean[k] = (float*)take(4 * d.npx);
        }
        for (int k = 0; k < 2; k++) P.tri[k] = (int32_t*)take(12 * (size_t)d.tricap);
        for (int k = 0; k < 2; k++) P.trirec[k] = (float*)take(4 * ELAS_TRIREC * (size_t)d.tricap);
        for (int k = 0; k < 2; k++) P.prior[k] = (int32_t*)take(4 * d.npx);
        P.ntx = (s.W + ELAS_RASTER_T - 1) / ELAS_RASTER_T;
        P.nty = (s.H + ELAS_RASTER_T - 1) / ELAS_RASTER_T;
        for (int k = 0; k < 2; k++) P.draw[k] = (float*)take(4 * d.npx);
        if (!s.dev_out) { P.dout[0] = (float*)take(4 * d.npx); P.dout[1] = (float*)take(4 * d.npx); }
        else { P.dout[0] = s.D1; P.dout[1] = s.D2; }
        sp_off[i] = so;
        so += 2 * 2 * d.ncan;
        tri_off[i] = to;
        to += 2 * 3 * (size_t)d.tricap;
    }
    auto tiles = [&](int i, int st) -> int {
        const ElasPairSpec& s = sp[i];
        const Dims& d = dd[i];
        const int gs = prm.grid_size, R = dense_rows(gs);
        switch (st) {
            case S_DESC: return cdiv(s.W, 64) * cdiv(s.H, 16) * 2;
            case S_SUPPORT: return d.ch * cdiv(d.cw, 64);
            case S_INCON: return cdiv(d.cw, 16) * cdiv(d.ch, 16);
            case S_REDUN_H: return cdiv(d.ch, 64);
            case S_REDUN_V: return cdiv(d.cw, 64);
            case S_BLOCKS: return d.nblk;
            case S_GRID_CELLS: return cdiv(2 * (long)d.ncell, 256);
            case S_GRID_FILL: return cdiv(2 * (long)d.ncell * d.words, 256);
            case S_PLANE: return cdiv((long)pairs[i].ntri[0] + pairs[i].ntri[1], 256);
            case S_RASTER: return 2 * pairs[i].ntx * pairs[i].nty;
            case S_DENSE: return 2 * d.gh * cdiv(gs, R) * cdiv(d.gw, ELAS_DENSE_WAVES * cpw);
            case S_PIXELS: return cdiv((long)d.npx, 256);
        }
        return 0;
    };
    auto build_prefix = [&](int32_t* out) {
        for (int st = 0; st < S_NUM_STAGES; st++) {
            int32_t* pf = out + (size_t)st * (n + 1);
            pf[0] = 0;
            for (int i = 0; i < n; i++) pf[i + 1] = pf[i] + tiles(i, st);
        }
    };
    build_prefix(prefix.data());
    memcpy(hpairs, pairs.data(), sizeof(ElasPair) * n);
    memcpy(hprefix, prefix.data(), sizeof(int32_t) * prefix.size());
    HIPCHECK(hipMemcpyAsync(dpairs[0], hpairs, sizeof(ElasPair) * n, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(dprefix[0], hprefix, sizeof(int32_t) * prefix.size(), hipMemcpyHostToDevice, stream));
    for (int i = 0; i < n; i++)
        if (!sp[i].dev_in) {
            HIPCHECK(hipMemcpyAsync((void*)pairs[i].img[0], sp[i].left, dd[i].npx, hipMemcpyHostToDevice, stream));
            HIPCHECK(hipMemcpyAsync((void*)pairs[i].img[1], sp[i].right, dd[i].npx, hipMemcpyHostToDevice, stream));
        }
    auto launch = [&](int table, int st) {
        ElasLaunch L;
        L.pairs = dpairs[table];
        L.prefix = dprefix[table] + (size_t)st * (n + 1);
        L.npairs = n;
        L.total = (table ? hprefix_b : hprefix)[(size_t)st * (n + 1) + n];
        return L;
    };
    const int cap256 = std::max(1, num_cu) * 8, cap1024 = std::max(1, num_cu) * 2;
    auto g = [&](const ElasLaunch& L, int cap) { return std::min(L.total, cap); };

    // ---- phase A: descriptors, support points, filters, compaction ----------
    HIPCHECK(hipEventRecord(ev[0], stream));
    { ElasLaunch L = launch(0, S_DESC); if (L.total) elas_launch_descriptor(L, g(L, cap256), stream); }
    HIPCHECK(hipEventRecord(ev[1], stream));
    { ElasLaunch L = launch(0, S_SUPPORT);
      // LDS: rows v-2 and v+2 over 63*step + dmax + 5 (left) or + 2*dmax + 5 (right) columns
      const size_t lds_l = 2 * 16 * (size_t)(63 * prm.candidate_stepsize + maxd + 5);
      const size_t lds_r = 2 * 16 * (size_t)(63 * prm.candidate_stepsize + 2 * maxd + 5);
      if (L.total) { elas_launch_support(L, kp, false, g(L, cap256), lds_l, stream);
                     elas_launch_support(L, kp, true, g(L, cap256), lds_r, stream); } }
    HIPCHECK(hipEventRecord(ev[2], stream));
    { ElasLaunch L = launch(0, S_INCON); if (L.total) elas_launch_incon(L, kp, g(L, cap256), stream); }
    { ElasLaunch L = launch(0, S_REDUN_H); if (L.total) elas_launch_redun(L, false, L.total, stream); }
    { ElasLaunch L = launch(0, S_REDUN_V); if (L.total) elas_launch_redun(L, true, L.total, stream); }
    HIPCHECK(hipEventRecord(ev[3], stream));
    { ElasLaunch L = launch(0, S_BLOCKS);
      if (L.total) { elas_launch_blkcount(L, g(L, cap256), stream); elas_launch_compact(L, g(L, cap256), stream); } }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(ev[4], stream));
    HIPCHECK(hipMemcpyAsync(hnsp, dnsp, 4 * (size_t)n, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipEventRecord(ev[5], stream));
    // disparity grid on the second stream, overlapping the host Triangle step
    HIPCHECK(hipStreamWaitEvent(stream2, ev[4], 0));
    HIPCHECK(hipEventRecord(ev[11], stream2));
    { ElasLaunch L = launch(0, S_GRID_FILL); if (L.total) elas_launch_grid_fill(L, g(L, cap256), stream2); }
    { ElasLaunch L = launch(0, S_BLOCKS); if (L.total) elas_launch_grid_scatter(L, g(L, cap256), stream2); }
    { ElasLaunch L = launch(0, S_GRID_CELLS);
      if (L.total) { elas_launch_grid_dilate(L, g(L, cap256), stream2);
                     if (debug) elas_launch_grid_list(L, g(L, cap256), stream2); } }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(ev[12], stream2));

    // ---- host: support points down, Triangle, triangles up ------------------
    HIPCHECK(hipEventSynchronize(ev[5]));
    const auto t_tri0 = clk::now();
    for (int i = 0; i < n; i++) {
        nsp[i] = dd[i].nblk > 0 ? hnsp[i] : 0;
        if (nsp[i] > 0) {
            HIPCHECK(hipMemcpyAsync(hsp + sp_off[i], pairs[i].sp[0], 8 * (size_t)nsp[i], hipMemcpyDeviceToHost, stream));
            HIPCHECK(hipMemcpyAsync(hsp + sp_off[i] + 2 * dd[i].ncan, pairs[i].sp[1], 8 * (size_t)nsp[i],
                                    hipMemcpyDeviceToHost, stream));
        }
    }
    HIPCHECK(hipStreamSynchronize(stream));
    std::vector<std::vector<int32_t>> tri_out(2 * (size_t)n), toff(2 * (size_t)n), tlist(2 * (size_t)n);
    std::vector<ElasTriJob> jobs;
    std::vector<int> job_of(2 * (size_t)n, -1);
    for (int i = 0; i < n; i++) {
        if (nsp[i] < 3) { status[i] = ELAS_ERR_TOO_FEW_POINTS; continue; }
        for (int k = 0; k < 2; k++) {
            job_of[2 * i + k] = (int)jobs.size();
            ElasTriJob j;
            j.pts = hsp + sp_off[i] + (size_t)k * 2 * dd[i].ncan;
            j.n = nsp[i];
            j.out = &tri_out[2 * i + k];
            j.result = 0;
            j.W = sp[i].W;
            j.H = sp[i].H;
            j.tile_off = &toff[2 * i + k];
            j.tile_list = &tlist[2 * i + k];
            jobs.push_back(j);
        }
    }
    elas_triangulate_jobs(jobs.data(), (int)jobs.size(), tri_workers);
    for (int i = 0; i < n; i++) {
        int nt[2] = {0, 0};
        if (status[i] == ELAS_OK) {
            for (int k = 0; k < 2; k++) {
                const int r = jobs[job_of[2 * i + k]].result;
                nt[k] = r < 0 ? 0 : r;
                if (r < 0) status[i] = ELAS_ERR_TOO_FEW_POINTS;
            }
            if (nt[0] > dd[i].tricap || nt[1] > dd[i].tricap)
                return fail(ELAS_ERR_HIP, "triangle count exceeds capacity");
        }
        if (status[i] != ELAS_OK) {
            nt[0] = nt[1] = 0;
            // a pair without a mesh rasterises nothing: every tile's list is empty
            for (int k = 0; k < 2; k++)
                if (job_of[2 * i + k] >= 0) {
                    std::fill(toff[2 * i + k].begin(), toff[2 * i + k].end(), 0);
                    tlist[2 * i + k].clear();
                }
        }
        ntri0[i] = nt[0];
        ntri1[i] = nt[1];
        pairs[i].ntri[0] = nt[0];
        pairs[i].ntri[1] = nt[1];
        for (int k = 0; k < 2; k++)
            if (nt[k] > 0) {
                int32_t* dst = htri + tri_off[i] + (size_t)k * 3 * dd[i].tricap;
                memcpy(dst, tri_out[2 * i + k].data(), 12 * (size_t)nt[k]);
                HIPCHECK(hipMemcpyAsync((void*)pairs[i].tri[k], dst, 12 * (size_t)nt[k], hipMemcpyHostToDevice, stream));
            }
    }
    // raster bins: [toff0 | toff1 | tlist0 | tlist1] per pair in one upload
    size_t bl_ints = 0;
    std::vector<size_t> bl_off(n);
    for (int i = 0; i < n; i++) {
        const size_t nt = (size_t)pairs[i].ntx * pairs[i].nty + 1;
        bl_off[i] = bl_ints;
        bl_ints += 2 * nt;
        for (int k = 0; k < 2; k++) bl_ints += (job_of[2 * i + k] >= 0) ? tlist[2 * i + k].size() : 0;
        bl_ints = (bl_ints + 63) & ~(size_t)63;
    }
    if (4 * bl_ints > blist_bytes) {
        if (blist) (void)hipFree(blist);
        blist = nullptr;
        blist_bytes = 0;
        if (hipMalloc(&blist, 4 * bl_ints + 4 * bl_ints / 4 + 256) != hipSuccess) {
            blist = nullptr;
            (void)hipGetLastError();
            return fail(ELAS_ERR_OUT_OF_MEMORY, "raster bin allocation failed");
        }
        blist_bytes = 4 * bl_ints + 4 * bl_ints / 4 + 256;
    }
    if (4 * bl_ints > blist_pin_bytes) {
        if (blist_pin) (void)hipHostFree(blist_pin);
        blist_pin = nullptr;
        blist_pin_bytes = 0;
        if (hipHostMalloc(&blist_pin, 4 * bl_ints + 4 * bl_ints / 4 + 256, hipHostMallocDefault) != hipSuccess) {
            blist_pin = nullptr;
            (void)hipGetLastError();
            return fail(ELAS_ERR_OUT_OF_MEMORY, "raster bin staging allocation failed");
        }
        blist_pin_bytes = 4 * bl_ints + 4 * bl_ints / 4 + 256;
    }
    {
        int32_t* hb = (int32_t*)blist_pin;
        int32_t* db = (int32_t*)blist;
        for (int i = 0; i < n; i++) {
            const size_t nt = (size_t)pairs[i].ntx * pairs[i].nty + 1;
            size_t o = bl_off[i];
            for (int k = 0; k < 2; k++) {
                int32_t* dst = hb + o;
                if (job_of[2 * i + k] >= 0) memcpy(dst, toff[2 * i + k].data(), 4 * nt);
                else memset(dst, 0, 4 * nt);
                pairs[i].toff[k] = db + o;
                o += nt;
            }
            for (int k = 0; k < 2; k++) {
                const size_t nl = (job_of[2 * i + k] >= 0) ? tlist[2 * i + k].size() : 0;
                if (nl) memcpy(hb + o, tlist[2 * i + k].data(), 4 * nl);
                pairs[i].tlist[k] = db + o;
                o += nl;
            }
        }
        if (bl_ints) HIPCHECK(hipMemcpyAsync(blist, blist_pin, 4 * bl_ints, hipMemcpyHostToDevice, stream));
    }
    build_prefix(prefix.data());
    memcpy(hpairs_b, pairs.data(), sizeof(ElasPair) * n);
    memcpy(hprefix_b, prefix.data(), sizeof(int32_t) * prefix.size());
    HIPCHECK(hipMemcpyAsync(dpairs[1], hpairs_b, sizeof(ElasPair) * n, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(dprefix[1], hprefix_b, sizeof(int32_t) * prefix.size(), hipMemcpyHostToDevice, stream));
    const auto t_tri1 = clk::now();
    HIPCHECK(hipStreamWaitEvent(stream, ev[12], 0));

    // ---- phase B: plane prior, dense matching, left/right check -------------
    HIPCHECK(hipEventRecord(ev[6], stream));
    { ElasLaunch L = launch(1, S_PLANE); if (L.total) elas_launch_plane(L, g(L, cap256), stream); }
    { ElasLaunch L = launch(1, S_RASTER); if (L.total) elas_launch_raster(L, g(L, cap256), stream); }
    HIPCHECK(hipEventRecord(ev[7], stream));
    {
        ElasLaunch L = launch(1, S_DENSE);
        if (L.total) elas_launch_dense(L, kpl, g(L, cap1024), dense_lds(prm.grid_size, maxd, cpw), stream);
    }
    HIPCHECK(hipEventRecord(ev[8], stream));
    { ElasLaunch L = launch(1, S_PIXELS); if (L.total) elas_launch_lr(L, kp, g(L, cap256), stream); }
    HIPCHECK(hipEventRecord(ev[9], stream));
    HIPCHECK(hipEventRecord(ev[10], stream));
    if (debug) { ElasLaunch L = launch(1, S_PIXELS); if (L.total) elas_launch_mean(L, g(L, cap256), stream); }
    HIPCHECK(hipGetLastError());
    for (int i = 0; i < n; i++)
        if (!sp[i].dev_out) {
            HIPCHECK(hipMemcpyAsync(sp[i].D1, pairs[i].dout[0], 4 * dd[i].npx, hipMemcpyDeviceToHost, stream));
            HIPCHECK(hipMemcpyAsync(sp[i].D2, pairs[i].dout[1], 4 * dd[i].npx, hipMemcpyDeviceToHost, stream));
        }
    HIPCHECK(hipStreamSynchronize(stream));
    const auto t_end = clk::now();
    auto el = [&](int a, int b) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[a], ev[b]);
        return ms;
    };
    times[ELAS_T_DESCRIPTOR] = el(0, 1);
    times[ELAS_T_SUPPORT] = el(1, 2);
    times[ELAS_T_FILTER] = el(2, 3);
    times[ELAS_T_COMPACT] = el(3, 4);
    times[ELAS_T_GRID] = el(11, 12);
    times[ELAS_T_TRIANGLE] = std::chrono::duration<float, std::milli>(t_tri1 - t_tri0).count();
    times[ELAS_T_PLANE] = el(6, 7);
    times[ELAS_T_DENSE] = el(7, 8);
    times[ELAS_T_LR] = el(8, 9);
    times[ELAS_T_POST] = el(9, 10);
    times[ELAS_T_TOTAL] = std::chrono::duration<float, std::milli>(t_end - t_start).count();
    int ret = ELAS_OK;
    for (int i = 0; i < n; i++) {
        if (statuses) statuses[i] = status[i];
        if (status[i] != ELAS_OK && status[i] != ELAS_ERR_TOO_FEW_POINTS) ret = status[i];
    }
    if (n == 1 && status[0] != ELAS_OK) {
        ret = status[0];
        err = "fewer than 3 support points (or a degenerate mesh): D1/D2 are all -1";
    }
    return ret;
}

int elas_ctx::fetch(int stage, void* dst, size_t bytes, size_t* stage_bytes) {
    if (pairs.empty()) return fail(ELAS_ERR_INVALID_ARGUMENT, "no process call yet");
    const ElasPair& P = pairs[0];
    const size_t npx = (size_t)P.W * P.H, ncan = (size_t)P.cw * P.ch;
    const void* src = nullptr;
    size_t sz = 0;
    switch (stage) {
        case ELAS_STAGE_DESC0: src = P.desc[0]; sz = 16 * npx; break;
        case ELAS_STAGE_DESC1: src = P.desc[1]; sz = 16 * npx; break;
        case ELAS_STAGE_CAN_LEFT: src = P.can[0]; sz = 4 * ncan; break;
        case ELAS_STAGE_CAN_LR: src = P.can[1]; sz = 4 * ncan; break;
        case ELAS_STAGE_CAN_INCON: src = P.can[2]; sz = 4 * ncan; break;
        case ELAS_STAGE_CAN_FINAL: src = P.can[3]; sz = 4 * ncan; break;
        case ELAS_STAGE_SP0: src = P.sp[0]; sz = 8 * (size_t)nsp[0]; break;
        case ELAS_STAGE_SP1: src = P.sp[1]; sz = 8 * (size_t)nsp[0]; break;
        case ELAS_STAGE_GRID0: src = P.glist[0]; sz = 4 * (size_t)P.gw * P.gh * (P.dmax + 2); break;
        case ELAS_STAGE_GRID1: src = P.glist[1]; sz = 4 * (size_t)P.gw * P.gh * (P.dmax + 2); break;
        case ELAS_STAGE_TRI0: src = P.tri[0]; sz = 12 * (size_t)ntri0[0]; break;
        case ELAS_STAGE_TRI1: src = P.tri[1]; sz = 12 * (size_t)ntri1[0]; break;
        case ELAS_STAGE_MEAN0: src = P.mean[0]; sz = 4 * npx; break;
        case ELAS_STAGE_MEAN1: src = P.mean[1]; sz = 4 * npx; break;
        case ELAS_STAGE_D1_RAW: src = P.draw[0]; sz = 4 * npx; break;
        case ELAS_STAGE_D2_RAW: src = P.draw[1]; sz = 4 * npx; break;
        case ELAS_STAGE_D1: src = P.dout[0]; sz = 4 * npx; break;
        case ELAS_STAGE_D2: src = P.dout[1]; sz = 4 * npx; break;
        default: return fail(ELAS_ERR_INVALID_ARGUMENT, "unknown stage");
    }
    if (stage_bytes) *stage_bytes = sz;
    if (!dst || bytes == 0 || sz == 0) return ELAS_OK;
    if (!src) return fail(ELAS_ERR_INVALID_ARGUMENT, "stage not kept (enable debug before the call)");
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(hipMemcpy(dst, src, std::min(bytes, sz), hipMemcpyDeviceToHost));
    return ELAS_OK;
}
