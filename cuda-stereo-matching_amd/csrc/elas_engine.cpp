// elas_engine.cpp -- host orchestration of the ELAS path on one gfx950 GPU.
//
// Stage order, buffer initial values and the host Delaunay excursion follow
// Elas::process (src/Elas.cu:19-120).  Differences of structure, none of
// arithmetic:
//   * every stage is one launch over a whole chunk of pairs (a single process
//     call is a chunk of one);
//   * device buffers live in per-slot arenas that grow on demand and are
//     reused across calls (the reference allocates ~20 buffers per call and
//     leaks its scan buffers, Elas.cu:349-371);
//   * the disparity grid (Elas.cu:65) runs while the host triangulates (on a
//     second stream, or after phase A when phase B waits behind its gate),
//     the left and right meshes are triangulated in parallel by independent
//     Triangle instances, and a batch keeps several chunks in flight so that
//     the GPU phases of one overlap the host step of another;
//   * a single-pair call queues phase B while the host triangulates, behind a
//     gate kernel that the host opens once the meshes are written, so no
//     launch cost follows the host step;
//   * buffers the reference initialises with thrust::fill (-1 for D1/D2 and
//     the plane map, 0 for the grid, Elas.cu:94-95,415-420,488) are written in
//     full by the kernels themselves, or cleared by fill kernels.
#include "elas_engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include <immintrin.h>

#include "elas_launch.h"
#include "elas_triangle.h"

#define HIPCHECK(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return fail(ELAS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

using clk = std::chrono::steady_clock;

// Waits for an event by polling for up to 2 ms, then blocking: a synchronous
// call waits here for GPU work of well under a millisecond, and a blocking
// wait would add the wake-up of a sleeping thread to it.
hipError_t poll_event(hipEvent_t e) {
    const auto t0 = clk::now();
    for (;;) {
        const hipError_t r = hipEventQuery(e);
        if (r != hipErrorNotReady) return r;
        if (clk::now() - t0 > std::chrono::milliseconds(2)) return hipEventSynchronize(e);
        _mm_pause();
    }
}

size_t al(size_t x) { return (x + 255) & ~size_t(255); }

int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

int dense_rows(int gs) { return std::max(1, std::min(64 / gs, gs)); }

// LDS of a dense unit: two rings (desc0, desc1) of R rows, each row the
// smallest power of two >= waves*cpw*gs + dmax slots of 16 B
size_t dense_lds(int gs, int dmax, int cpw) {
    size_t rw = 1;
    while (rw < (size_t)(ELAS_DENSE_WAVES * cpw * gs + dmax)) rw <<= 1;
    return 2 * (size_t)dense_rows(gs) * rw * 16;
}

// steps of a band and work units (band x chunk) of one pair's dense launch
int dense_segs(int gw, int cpw) { return cdiv(gw, ELAS_DENSE_WAVES * cpw); }
int dense_units(const ElasDims& d, int gs, int cpw, int spc) {
    return d.gh * cdiv(gs, dense_rows(gs)) * cdiv(dense_segs(d.gw, cpw), spc);
}

// descriptor strip heights and the strips of one pair (both images)
const int kDescRows[3] = {6, 12, 18};
int desc_units(int W, int H, int sh) { return 2 * cdiv(W, ELAS_DESC_SW) * cdiv(H, sh); }

// Grid of a dense launch: the units rounded up to a multiple of 8 (so that a
// workgroup's XCD is its index mod 8 in every grid-stride step), at most 4
// workgroups of 512 lanes per CU.
int dense_grid(int units, int num_cu) {
    const int cap = std::max(1, num_cu) * 4;
    return std::min((units + 7) & ~7, (cap + 7) & ~7);
}

// LDS of a support workgroup: rows v-2 and v+2 over 63*step + dmax + 5
// (left pass) or 63*step + 2*dmax + 5 (right pass) columns
size_t support_lds(int step, int dmax, bool right) {
    return 2 * 16 * (size_t)(63 * step + (right ? 2 : 1) * dmax + 5);
}

bool post_scratch(const elas_params& p) { return p.speckle_size > 0 || p.filter_adaptive_mean || p.filter_median; }

// device bytes of one pair in a slot arena
size_t pair_arena(const ElasPairSpec& s, const ElasDims& d, bool debug, bool post) {
    size_t b = 0;
    if (!s.dev_in) b += 2 * al(d.npx);
    b += 2 * al(16 * d.npx);
    b += 4 * al(4 * d.ncan) + al(4 * (size_t)d.nblk) + 2 * al(8 * d.nspcap);
    b += 2 * al(4 * d.ncell * d.words) + al(4 * d.ncell * (size_t)elas_cand_stride(s.dmax));
    if (debug) b += 2 * al(4 * d.ncell * (size_t)(s.dmax + 2)) + 2 * al(4 * d.npx);
    b += 2 * al(4 * ELAS_TRIREC * (size_t)d.tricap);
    b += 2 * al(4 * (size_t)d.nst) + 2 * al(16 * (size_t)d.nst * d.scap);  // culling lists
    b += 2 * al(4 * d.npx) + 2 * al(4 * d.npx);  // prior, draw
    if (!s.dev_out) b += 2 * al(4 * d.npx);
    if (post) b += 2 * al(8 * d.npx);  // post-filter scratch
    return b;
}

// one device table: pairs, per-stage tile prefixes, support point counts
size_t table_one(int n) {
    return al(sizeof(ElasPair) * n) + al(sizeof(int32_t) * S_NUM_STAGES * (n + 1)) + al(4 * (size_t)n);
}

}  // namespace

ElasDims elas_dims(const elas_params& p, int W, int H, int dmax) {
    ElasDims d;
    d.cw = W / p.candidate_stepsize;
    d.ch = H / p.candidate_stepsize;
    d.gw = (int32_t)ceil((float)W / (float)p.grid_size);  // Elas.cu:411-412
    d.gh = (int32_t)ceil((float)H / (float)p.grid_size);
    d.words = (dmax + 32) / 32;
    d.npx = (size_t)W * H;
    d.ncan = (size_t)d.cw * d.ch;
    d.ncell = (size_t)d.gw * d.gh;
    d.nblk = (int32_t)((d.ncan + 255) / 256);
    d.nspcap = d.ncan + 8;  // the lattice's points and the corner points of add_corners
    d.tricap = (int32_t)(2 * d.nspcap + 8);
    // raster culling super-tiles: single raster tiles, or squares of 2, 4, ..
    // tiles a side where an image has more than ELAS_ST_MAX tiles
    const int ntx = cdiv(W, ELAS_RT_W), nty = cdiv(H, ELAS_RT_H);
    d.sshift = 0;
    while ((long)cdiv(ntx, 1 << d.sshift) * cdiv(nty, 1 << d.sshift) > ELAS_ST_MAX) d.sshift++;
    d.nsx = cdiv(ntx, 1 << d.sshift);
    d.nst = d.nsx * cdiv(nty, 1 << d.sshift);
    d.scap = std::min(cdiv(d.tricap, 64), ELAS_SCAP_MAX);
    return d;
}

size_t elas_pair_device_bytes(const elas_params& p, const ElasPairSpec& s, bool debug) {
    return pair_arena(s, elas_dims(p, s.W, s.H, s.dmax), debug, post_scratch(p));
}

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
    if (p->speckle_size < 0 || p->ipol_gap_width < 0 || !(p->speckle_sim_threshold >= 0.0f))
        return bad("speckle_size, ipol_gap_width and speckle_sim_threshold must be >= 0");
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
    if (pmin < -ELAS_MAX_KEY_BIAS) return bad("prior too strong (beta too small): -P(k) must be <= 16384");
    kp->key_bias = -pmin;
    kp->all_p_negative = allneg ? 1 : 0;
    kp->speckle_sim = p->speckle_sim_threshold;
    kp->speckle_size = p->speckle_size;
    kp->ipol_gap = p->ipol_gap_width;
    kp->adaptive_mean = p->filter_adaptive_mean ? 1 : 0;
    kp->median = p->filter_median ? 1 : 0;
    kp->post_images = p->postprocess_only_left ? 1 : 2;
    kp->add_corners = p->add_corners ? 1 : 0;
    if (p->subsampling) {
        // src/Elas.h:36-38: D1/D2 of W/2 x H/2; the reference gives no
        // semantics beyond that comment (which pixels, and the left/right
        // check at half resolution), so it is refused rather than guessed
        *err = "subsampling is not supported (the reference declares it without semantics)";
        return ELAS_ERR_UNSUPPORTED;
    }
    return ELAS_OK;
}

// ---- slot lifetime ---------------------------------------------------------

int elas_ctx::slot_init(ElasSlot& S) {
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&S.stream2, hipStreamNonBlocking));
    for (auto& e : S.ev) HIPCHECK(hipEventCreate(&e));
    HIPCHECK(hipHostMalloc((void**)&S.gate, 256, hipHostMallocMapped | hipHostMallocCoherent));
    memset(S.gate, 0, 256);
    HIPCHECK(hipHostGetDevicePointer((void**)&S.gate_dev, S.gate, 0));
    return ELAS_OK;
}

// Sets the phase B release flag if a gate is waiting on it (every path out
// of a gated chunk goes through here or through slot_phase_b).
void elas_ctx::gate_release(ElasSlot& S) {
    if (!S.gate_armed) return;
    __atomic_store_n(&S.gate[0], 1u, __ATOMIC_RELEASE);
    S.gate_armed = false;
}

void elas_ctx::slot_free(ElasSlot& S) {
    (void)hipSetDevice(device);
    gate_release(S);
    if (S.stream) (void)hipStreamSynchronize(S.stream);
    if (S.stream2) (void)hipStreamSynchronize(S.stream2);
    slot_drop_graphs(S);
    if (S.ticket) elas_tri_release(S.ticket);
    S.ticket = nullptr;
    for (auto& e : S.ev)
        if (e) (void)hipEventDestroy(e);
    if (S.arena && !S.external) (void)hipFree(S.arena);
    if (S.table && !S.external) (void)hipFree(S.table);
    if (S.pin) (void)hipHostFree(S.pin);
    if (S.tune) (void)hipFree(S.tune);
    if (S.stream) (void)hipStreamDestroy(S.stream);
    if (S.stream2) (void)hipStreamDestroy(S.stream2);
    if (S.gate) (void)hipHostFree(S.gate);
}

// Validates a chunk; returns the arena and pinned bytes it needs.
int elas_ctx::slot_need(const ElasPairSpec* sp, int n, size_t* arena_need, size_t* pin_need) {
    if (n <= 0 || !sp) return fail(ELAS_ERR_INVALID_ARGUMENT, "empty batch");
    size_t a = 0, psp = 0, ptri = 0;
    for (int i = 0; i < n; i++) {
        const ElasPairSpec& s = sp[i];
        if (s.W < 1 || s.H < 1 || s.W > 16384 || s.H > 16384 || (size_t)s.W * s.H > (size_t)1 << 27)
            return fail(ELAS_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": dims out of range");
        if (!s.left || !s.right || !s.D1 || !s.D2)
            return fail(ELAS_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": null buffer");
        if (s.dmax < std::max(prm.disp_min, 0) || s.dmax > ELAS_MAX_DISP)
            return fail(ELAS_ERR_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": disp_max out of range");
        if (dense_lds(prm.grid_size, s.dmax, 1) > max_lds)
            return fail(ELAS_ERR_UNSUPPORTED, "grid_size/disp_max exceed the dense LDS rings");
        if (support_lds(prm.candidate_stepsize, s.dmax, true) > 65536)
            return fail(ELAS_ERR_UNSUPPORTED, "candidate_stepsize/disp_max exceed the support LDS rows");
        if (prm.ipol_gap_width > 0 && std::max(s.W, s.H) > ELAS_GAP_MAX_LINE)
            return fail(ELAS_ERR_UNSUPPORTED, "gap interpolation supports lines up to 15872 pixels");
        const ElasDims d = elas_dims(prm, s.W, s.H, s.dmax);
        a += pair_arena(s, d, debug, post_scratch(prm));
        psp += 2 * 8 * d.nspcap;
        ptri += 2 * 12 * (size_t)d.tricap;
    }
    *arena_need = a;
    *pin_need = 3 * table_one(n) + al(psp) + al(ptri);
    return ELAS_OK;
}

// Grows a slot's buffers; the slot must be idle.
int elas_ctx::slot_reserve(ElasSlot& S, size_t arena_need, size_t pin_need, int n) {
    HIPCHECK(hipSetDevice(device));
    if (arena_need > S.arena_bytes) {
        slot_drop_graphs(S);
        if (S.arena) (void)hipFree(S.arena);
        S.arena = nullptr;
        S.arena_bytes = 0;
        const size_t want = arena_need + arena_need / 4;
        if (hipMalloc(&S.arena, want) != hipSuccess) {
            S.arena = nullptr;
            (void)hipGetLastError();
            return fail(ELAS_ERR_OUT_OF_MEMORY, "arena allocation of " + std::to_string(want) + " bytes failed");
        }
        S.arena_bytes = want;
    }
    const size_t tab = 2 * table_one(n);
    if (tab > S.table_bytes) {
        slot_drop_graphs(S);
        if (S.table) (void)hipFree(S.table);
        S.table = nullptr;
        S.table_bytes = 0;
        S.tab0_up.clear();
        if (hipMalloc(&S.table, tab + tab / 4) != hipSuccess) {
            S.table = nullptr;
            (void)hipGetLastError();
            return fail(ELAS_ERR_OUT_OF_MEMORY, "table allocation failed");
        }
        S.table_bytes = tab + tab / 4;
    }
    if (pin_need > S.pin_bytes) {
        slot_drop_graphs(S);
        if (S.pin) (void)hipHostFree(S.pin);
        S.pin = nullptr;
        S.pin_dev = nullptr;
        S.pin_bytes = 0;
        // mapped and coherent: the compaction kernel writes the support
        // points and their count straight into it for the host step
        if (hipHostMalloc(&S.pin, pin_need + pin_need / 4, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
            S.pin = nullptr;
            (void)hipGetLastError();
            return fail(ELAS_ERR_OUT_OF_MEMORY, "pinned staging allocation failed");
        }
        if (hipHostGetDevicePointer((void**)&S.pin_dev, S.pin, 0) != hipSuccess) {
            (void)hipHostFree(S.pin);
            S.pin = nullptr;
            S.pin_dev = nullptr;
            (void)hipGetLastError();
            return fail(ELAS_ERR_OUT_OF_MEMORY, "pinned staging is not mapped for the device");
        }
        S.pin_bytes = pin_need + pin_need / 4;
    }
    return ELAS_OK;
}

// ---- chunk layout ------------------------------------------------------------

int elas_ctx::slot_prepare(ElasSlot& S, const ElasPairSpec* sp, int n) {
    S.t_start = clk::now();
    S.gated = false;
    S.gate_used = false;
    size_t arena_need = 0, pin_need = 0;
    int st = slot_need(sp, n, &arena_need, &pin_need);
    if (st != ELAS_OK) return st;
    if (S.external) {
        // buffers are in place (run_fixed checked that the pair fits)
        if (arena_need > S.arena_bytes || pin_need > S.pin_bytes || 2 * table_one(n) > S.table_bytes)
            return fail(ELAS_ERR_UNSUPPORTED, "pair does not fit the scratch of the no-allocation entry point");
    } else {
        st = slot_reserve(S, arena_need, pin_need, n);
        if (st != ELAS_OK) return st;
    }
    S.n = n;
    S.specs.assign(sp, sp + n);
    S.dd.resize(n);
    S.base.resize(n);
    S.sp_off.resize(n);
    S.tri_off.resize(n);
    S.maxd = 0;
    size_t a = 0, so = 0, to = 0, psp = 0, ptri = 0;
    for (int i = 0; i < n; i++) {
        S.dd[i] = elas_dims(prm, sp[i].W, sp[i].H, sp[i].dmax);
        S.base[i] = a;
        a += pair_arena(sp[i], S.dd[i], debug, post_scratch(prm));
        S.sp_off[i] = so;
        so += 2 * 2 * S.dd[i].nspcap;
        S.tri_off[i] = to;
        to += 2 * 3 * (size_t)S.dd[i].tricap;
        psp += 2 * 8 * S.dd[i].nspcap;
        ptri += 2 * 12 * (size_t)S.dd[i].tricap;
        S.maxd = std::max(S.maxd, sp[i].dmax);
    }
    S.kpl = kp;
    S.pin_sp = psp;
    S.pin_tri = ptri;
    // device tables (phase A / phase B) and their pinned mirrors
    S.tab_pairs = al(sizeof(ElasPair) * n);
    S.tab_prefix = al(sizeof(int32_t) * S_NUM_STAGES * (n + 1));
    S.tab_one = S.tab_pairs + S.tab_prefix + al(4 * (size_t)n);
    char* A = (char*)S.arena;
    char* T = (char*)S.table;
    char* PIN = (char*)S.pin;
    S.dpairs[0] = (ElasPair*)T;
    S.dpairs[1] = (ElasPair*)(T + S.tab_one);
    S.dprefix[0] = (int32_t*)(T + S.tab_pairs);
    S.dprefix[1] = (int32_t*)(T + S.tab_one + S.tab_pairs);
    S.dnsp = (int32_t*)(T + S.tab_pairs + S.tab_prefix);
    S.hpairs[0] = (ElasPair*)PIN;
    S.hprefix[0] = (int32_t*)(PIN + S.tab_pairs);
    S.hnsp = (int32_t*)(PIN + S.tab_pairs + S.tab_prefix);
    S.hpairs[1] = (ElasPair*)(PIN + S.tab_one);
    S.hprefix[1] = (int32_t*)(PIN + S.tab_one + S.tab_pairs);
    S.hpairs[2] = (ElasPair*)(PIN + 2 * S.tab_one);
    S.hprefix[2] = (int32_t*)(PIN + 2 * S.tab_one + S.tab_pairs);
    S.hsp = (float*)(PIN + 3 * S.tab_one);
    S.htri = (int32_t*)(PIN + 3 * S.tab_one + al(psp));

    S.pairs.assign(n, ElasPair());
    S.dense_lds[0] = S.dense_lds[1] = 0;
    S.plane_lds = 0;
    S.prefix.assign((size_t)S_NUM_STAGES * (n + 1), 0);
    S.nsp.assign(n, 0);
    S.ntri0.assign(n, 0);
    S.ntri1.assign(n, 0);
    S.status.assign(n, ELAS_OK);
    for (int i = 0; i < n; i++) {
        const ElasPairSpec& s = sp[i];
        const ElasDims& d = S.dd[i];
        ElasPair& P = S.pairs[i];
        memset(&P, 0, sizeof(P));
        P.W = s.W; P.H = s.H; P.dmin = prm.disp_min; P.dmax = s.dmax;
        P.cw = d.cw; P.ch = d.ch; P.gw = d.gw; P.gh = d.gh; P.words = d.words;
        P.cstride = elas_cand_stride(s.dmax);
        P.step = prm.candidate_stepsize; P.gsize = prm.grid_size; P.nblk = d.nblk;
        const ElasTiles dt = tiles(d, s.W, s.H, s.dmax);
        P.dcpw = dt.cpw;
        P.dspc = dt.spc;
        P.dsh = dt.dsh;
        S.dense_lds[dt.cpw - 1] = std::max(S.dense_lds[dt.cpw - 1], dense_lds(prm.grid_size, s.dmax, dt.cpw));
        char* c = A + S.base[i];
        auto take = [&](size_t bytes) { char* r = c; c += al(bytes); return r; };
        if (!s.dev_in) { P.img[0] = (uint8_t*)take(d.npx); P.img[1] = (uint8_t*)take(d.npx); }
        else { P.img[0] = s.left; P.img[1] = s.right; }
        for (int k = 0; k < 2; k++) P.desc[k] = (uint8_t*)take(16 * d.npx);
        for (int k = 0; k < 4; k++) P.can[k] = (int32_t*)take(4 * d.ncan);
        P.blkcnt = (int32_t*)take(4 * (size_t)d.nblk);
        for (int k = 0; k < 2; k++) P.sp[k] = (float*)take(8 * d.nspcap);
        P.nsp = S.dnsp + i;
        // device addresses in the mapped pinned staging: the support points
        // and their count again (read by the host step without a copy), and
        // the meshes (written there by the host step's Triangle workers, read
        // by the plane kernel without a copy; the right mesh tricap
        // triangles after the left one)
        auto hdev = [&](const void* h) { return S.pin_dev + ((const char*)h - (const char*)S.pin); };
        P.hsp[0] = (float*)hdev(S.hsp + S.sp_off[i]);
        P.hsp[1] = (float*)hdev(S.hsp + S.sp_off[i] + 2 * d.nspcap);
        P.hnsp = (int32_t*)hdev(S.hnsp + i);
        P.tri[0] = (const int32_t*)hdev(S.htri + S.tri_off[i]);
        P.tri[1] = P.tri[0] + 3 * (size_t)d.tricap;
        for (int k = 0; k < 2; k++) P.gflag[k] = (uint32_t*)take(4 * d.ncell * d.words);
        // both candidate lists in one block (the dense kernel reads them through one buffer resource)
        P.gcand[0] = (uint16_t*)take(4 * d.ncell * (size_t)elas_cand_stride(s.dmax));
        P.gcand[1] = P.gcand[0] + d.ncell * (size_t)elas_cand_stride(s.dmax);
        if (debug) {
            for (int k = 0; k < 2; k++) P.glist[k] = (int32_t*)take(4 * d.ncell * (size_t)(s.dmax + 2));
            for (int k = 0; k < 2; k++) P.mean[k] = (float*)take(4 * d.npx);
        }
        for (int k = 0; k < 2; k++) P.trirec[k] = (float*)take(4 * ELAS_TRIREC * (size_t)d.tricap);
        for (int k = 0; k < 2; k++) P.scnt[k] = (int32_t*)take(4 * (size_t)d.nst);
        for (int k = 0; k < 2; k++) P.slist[k] = (int32_t*)take(16 * (size_t)d.nst * d.scap);
        P.sshift = d.sshift; P.nsx = d.nsx; P.nst = d.nst; P.scap = tiny_lists ? 1 : d.scap;
        S.plane_lds = std::max(S.plane_lds, 8 * (size_t)d.nst);
        for (int k = 0; k < 2; k++) P.prior[k] = (int32_t*)take(4 * d.npx);
        for (int k = 0; k < 2; k++) P.draw[k] = (float*)take(4 * d.npx);
        if (!s.dev_out) { P.dout[0] = (float*)take(4 * d.npx); P.dout[1] = (float*)take(4 * d.npx); }
        else { P.dout[0] = s.D1; P.dout[1] = s.D2; }
        if (post_scratch(prm))
            for (int k = 0; k < 2; k++) P.pbuf[k] = (int32_t*)take(8 * d.npx);
    }
    return ELAS_OK;
}

namespace {

// workgroup tiles of stage st for one pair
int stage_tiles(const elas_params& prm, const ElasPairSpec& s, const ElasDims& d, const ElasPair& P, int st) {
    const int gs = prm.grid_size, R = dense_rows(gs);
    switch (st) {
        case S_DESC: return P.dsh == 18 ? desc_units(s.W, s.H, 18) : 0;
        case S_DESC12: return P.dsh == 12 ? desc_units(s.W, s.H, 12) : 0;
        case S_DESC6: return P.dsh == 6 ? desc_units(s.W, s.H, 6) : 0;
        case S_SUPPORT: return d.ch * cdiv(d.cw, 64);
        case S_INCON: return cdiv(d.cw, 16) * cdiv(d.ch, 16);
        case S_REDUN_H: return cdiv(d.ch, elas_redun_lines(d.cw));
        case S_REDUN_V: return cdiv(d.cw, elas_redun_lines(d.ch));
        case S_BLOCKS: return d.nblk;
        case S_GRID_CELLS: return cdiv(2 * (long)d.ncell, 256);
        case S_GRID_FILL: return cdiv(2 * (long)d.ncell * d.words, 256);
        case S_PLANE: return cdiv(P.ntri[0], 64) + cdiv(P.ntri[1], 64);
        case S_RTILES: return 2 * cdiv(s.W, ELAS_RT_W) * cdiv(s.H, ELAS_RT_H);
        case S_DENSE: return P.dcpw == 1 ? dense_units(d, gs, 1, P.dspc) : 0;
        case S_DENSE2: return P.dcpw == 2 ? dense_units(d, gs, 2, P.dspc) : 0;
        case S_PIXELS: return cdiv((long)d.npx, 256);
        case S_POST_TILES: return 2 * cdiv(s.W, ELAS_POST_TW) * cdiv(s.H, ELAS_POST_TH);
        case S_POST_ROWS: return 2 * s.H;
        case S_POST_COLS: return 2 * s.W;
        case S_PAIRS: return 1;
    }
    return 0;
}

void build_prefix(const elas_params& prm, ElasSlot& S) {
    const int n = S.n;
    for (int st = 0; st < S_NUM_STAGES; st++) {
        int32_t* pf = S.prefix.data() + (size_t)st * (n + 1);
        pf[0] = 0;
        for (int i = 0; i < n; i++) pf[i + 1] = pf[i] + stage_tiles(prm, S.specs[i], S.dd[i], S.pairs[i], st);
    }
}

ElasLaunch launch_of(const ElasSlot& S, int table, int st) {
    ElasLaunch L;
    L.pairs = S.dpairs[table];
    L.prefix = S.dprefix[table] + (size_t)st * (S.n + 1);
    L.npairs = S.n;
    L.total = S.hostpre[table][(size_t)st * (S.n + 1) + S.n];
    return L;
}

}  // namespace

// ---- phase A: descriptors, support points, filters, compaction, grid -------

// Phase A on the pipeline stream: descriptors .. compaction, which writes the
// support points and their count into the pinned staging as well (the table
// goes up before, in slot_phase_a).  With `timed`, stage events are recorded
// between the kernels (not possible inside a captured graph).
int elas_ctx::enqueue_a(ElasSlot& S, bool timed) {
    const int n = S.n;
    hipStream_t s1 = S.stream;
    const int cap256 = std::max(1, num_cu) * 8;
    auto g = [&](const ElasLaunch& L, int cap) { return std::min(L.total, cap); };
    auto ev = [&](int i) { return timed ? hipEventRecord(S.ev[i], s1) : hipSuccess; };
    HIPCHECK(ev(1));
    for (int sh : {18, 12, 6}) {
        const ElasLaunch L = launch_of(S, 0, sh == 18 ? S_DESC : sh == 12 ? S_DESC12 : S_DESC6);
        if (L.total) elas_launch_descriptor(L, sh, g(L, cap256), s1);
    }
    HIPCHECK(ev(2));
    { ElasLaunch L = launch_of(S, 0, S_SUPPORT);
      if (L.total) {
          elas_launch_support(L, kp, g(L, cap256), support_lds(prm.candidate_stepsize, S.maxd, true), s1);
      } }
    HIPCHECK(ev(3));
    { ElasLaunch L = launch_of(S, 0, S_INCON); if (L.total) elas_launch_incon(L, kp, g(L, cap256), s1); }
    {
        size_t lh = 4, lv = 4;
        for (int i = 0; i < n; i++) {
            lh = std::max(lh, (size_t)elas_redun_lds(S.dd[i].cw));
            lv = std::max(lv, (size_t)elas_redun_lds(S.dd[i].ch));
        }
        ElasLaunch L = launch_of(S, 0, S_REDUN_H);
        if (L.total) elas_launch_redun(L, false, g(L, cap256), lh, s1);
        L = launch_of(S, 0, S_REDUN_V);
        if (L.total) elas_launch_redun(L, true, g(L, cap256), lv, s1);
    }
    HIPCHECK(ev(4));
    { ElasLaunch L = launch_of(S, 0, S_BLOCKS);
      if (L.total) { elas_launch_blkcount(L, g(L, cap256), s1); elas_launch_compact(L, g(L, cap256), s1); } }
    if (kp.add_corners) {
        const ElasLaunch L = launch_of(S, 0, S_PAIRS);
        if (L.total) elas_launch_corners(L, g(L, cap256), s1);
    }
    HIPCHECK(hipGetLastError());
    return ELAS_OK;
}

int elas_ctx::slot_phase_a(ElasSlot& S) {
    const int n = S.n;
    HIPCHECK(hipSetDevice(device));
    build_prefix(prm, S);
    memcpy(S.hpairs[0], S.pairs.data(), sizeof(ElasPair) * n);
    memcpy(S.hprefix[0], S.prefix.data(), sizeof(int32_t) * S.prefix.size());
    S.hostpre[0] = S.prefix;
    hipStream_t s1 = S.stream, s2 = S.stream2;
    // pair records and stage prefixes lie back to back in both tables: one
    // copy, skipped when the device table already holds these bytes (calls
    // repeating a shape on the same buffers)
    {
        const size_t bytes = S.tab_pairs + sizeof(int32_t) * S.prefix.size();
        if (S.tab0_up.size() != bytes || memcmp(S.tab0_up.data(), S.hpairs[0], bytes) != 0) {
            HIPCHECK(hipMemcpyAsync(S.dpairs[0], S.hpairs[0], bytes, hipMemcpyHostToDevice, s1));
            S.tab0_up.assign((const char*)S.hpairs[0], (const char*)S.hpairs[0] + bytes);
        }
    }
    for (int i = 0; i < n; i++)
        if (!S.specs[i].dev_in) {
            HIPCHECK(hipMemcpyAsync((void*)S.pairs[i].img[0], S.specs[i].left, S.dd[i].npx, hipMemcpyHostToDevice, s1));
            HIPCHECK(hipMemcpyAsync((void*)S.pairs[i].img[1], S.specs[i].right, S.dd[i].npx, hipMemcpyHostToDevice, s1));
        }
    const int cap256 = std::max(1, num_cu) * 8;
    auto g = [&](const ElasLaunch& L, int cap) { return std::min(L.total, cap); };
    HIPCHECK(hipEventRecord(S.ev[0], s1));
    S.graphed = false;
    if (use_graphs && !S.gated) {
        hipGraphExec_t ex = nullptr;
        int r = slot_graph(S, 0, &ex);
        if (r != ELAS_OK) return r;
        HIPCHECK(hipGraphLaunch(ex, s1));
        S.graphed = true;
    } else {
        // a single-pair call launches phase A directly: a graph launch costs
        // the host about 20 us before the first kernel can start, a launch
        // about 5 us, and phase A's kernels mostly outlast a launch
        int r = enqueue_a(S, !use_graphs);
        if (r != ELAS_OK) return r;
        S.graphed = use_graphs;
    }
    HIPCHECK(hipEventRecord(S.ev[5], s1));
    // a single-pair call: while the device computes the support points, the
    // two workers that will take the meshes warm their order scratch (sized
    // by the slot's last call)
    if (S.gated && n == 1 && S.last_nsp >= 3 && tri_workers > 1) elas_tri_prewarm(S.last_nsp, S.last_nsp);
    // disparity grid on the second stream, overlapping the host Triangle
    // step; a gated chunk (one slot, phase B queued behind the gate) keeps it
    // on the first stream after phase A, where the device would only wait
    if (S.gated) s2 = s1;
    else HIPCHECK(hipStreamWaitEvent(s2, S.ev[5], 0));
    HIPCHECK(hipEventRecord(S.ev[11], s2));
    { ElasLaunch L = launch_of(S, 0, S_GRID_FILL); if (L.total) elas_launch_grid_fill(L, g(L, cap256), s2); }
    { ElasLaunch L = launch_of(S, 0, S_BLOCKS); if (L.total) elas_launch_grid_scatter(L, g(L, cap256), s2); }
    { ElasLaunch L = launch_of(S, 0, S_GRID_CELLS);
      if (L.total) { elas_launch_grid_dilate(L, g(L, cap256), s2);
                     if (debug) elas_launch_grid_list(L, g(L, cap256), s2); } }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(S.ev[12], s2));
    // phase B (on s1) needs the grid: the wait is queued now, while the GPU
    // runs phase A, rather than after the host step
    if (s2 != s1) HIPCHECK(hipStreamWaitEvent(s1, S.ev[12], 0));
    S.t_alaunch = clk::now();
    S.phase = SLOT_A;
    return ELAS_OK;
}

bool elas_ctx::slot_a_ready(ElasSlot& S) { return hipEventQuery(S.ev[5]) == hipSuccess; }

// ---- host step: support points read from pinned memory, Triangle jobs queued --

int elas_ctx::slot_start_tri(ElasSlot& S) {
    const int n = S.n;
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(poll_event(S.ev[5]));
    S.t_tri0 = clk::now();
    // the compaction kernel wrote the counts and support points into the
    // pinned staging; phase A is complete, so they are all there
    for (int i = 0; i < n; i++) S.nsp[i] = S.dd[i].nblk > 0 ? S.hnsp[i] : 0;
    if (n == 1) S.last_nsp = S.nsp[0];
    S.t_sp = clk::now();
    S.jobs.clear();
    S.job_of.assign(2 * (size_t)n, -1);
    for (int i = 0; i < n; i++) {
        // meshes of fewer than 65536 points cross PCIe as 16-bit corner indices
        S.pairs[i].tri16 = S.nsp[i] <= 65536 ? 1 : 0;
        if (S.nsp[i] < 3) { S.status[i] = ELAS_ERR_TOO_FEW_POINTS; continue; }
        for (int k = 0; k < 2; k++) {
            S.job_of[2 * i + k] = (int)S.jobs.size();
            ElasTriJob j;
            j.pts = S.hsp + S.sp_off[i] + (size_t)k * 2 * S.dd[i].nspcap;
            j.n = S.nsp[i];
            j.out = nullptr;
            j.dst = S.htri + S.tri_off[i] + (size_t)k * 3 * S.dd[i].tricap;
            j.cap = S.dd[i].tricap;
            j.dst16 = S.pairs[i].tri16 != 0;
            j.result = 0;
            S.jobs.push_back(j);
        }
    }
    S.t_sub = clk::now();
    if (tri_workers <= 1) {
        elas_triangulate_jobs(S.jobs.data(), (int)S.jobs.size(), 1);
        S.ticket = elas_tri_submit(nullptr, 0);
    } else {
        S.ticket = elas_tri_submit(S.jobs.data(), (int)S.jobs.size());
    }
    S.phase = SLOT_TRI;
    return S.gated ? slot_prelaunch_b(S) : ELAS_OK;
}

// Gated phase B, queued while the meshes are being computed: a placeholder
// table (no triangles; every pointer valid) goes to pinned memory, then the
// gate kernel and phase B are launched; slot_phase_b writes the real table
// and sets the flag.
int elas_ctx::slot_prelaunch_b(ElasSlot& S) {
    const int n = S.n;
    for (int i = 0; i < n; i++) {
        S.pairs[i].ntri[0] = S.pairs[i].ntri[1] = 0;
    }
    build_prefix(prm, S);
    // Nothing between the gate and its release may wait for the stream: no
    // graph capture or instantiation and no copy to pageable memory (the
    // result copies are queued after the release).  Phase B's kernels are
    // launched directly behind the gate (their launch cost is hidden by the
    // host step, and the device starts the first one sooner after the gate
    // than a graph's).
    memcpy(S.hpairs[1], S.pairs.data(), sizeof(ElasPair) * n);
    memcpy(S.hprefix[1], S.prefix.data(), sizeof(int32_t) * S.prefix.size());
    S.hostpre[1] = S.prefix;
    __atomic_store_n(&S.gate[0], 0u, __ATOMIC_RELEASE);
    S.gate[1] = 0u;
    hipStream_t s1 = S.stream;
    auto hdev = [&](const void* h) { return S.pin_dev + ((const char*)h - (const char*)S.pin); };
    const size_t bytes = S.tab_pairs + sizeof(int32_t) * S.prefix.size();
    S.gate_armed = true;
    S.gate_used = true;
    // released: the real table (hpairs[2], written by slot_phase_b before the
    // flag); timed out: this placeholder (hpairs[1], not written again while
    // the gate is armed, and with no triangles, so phase B reads no mesh)
    elas_launch_gate(S.gate_dev, S.gate_dev + 1, (const uint32_t*)hdev(S.hpairs[2]),
                     (const uint32_t*)hdev(S.hpairs[1]), (uint32_t*)S.dpairs[1], (int)((bytes + 3) / 4), gate_ticks,
                     s1);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(S.ev[6], s1));
    int r = enqueue_b(S, !use_graphs);
    if (r != ELAS_OK) return r;
    HIPCHECK(hipEventRecord(S.ev[10], s1));
    return ELAS_OK;
}

bool elas_ctx::slot_tri_ready(ElasSlot& S) { return elas_tri_done(S.ticket); }

// ---- phase B: plane prior, dense matching, left/right check ------------------

int elas_ctx::slot_phase_b(ElasSlot& S) {
    const int n = S.n;
    HIPCHECK(hipSetDevice(device));
    elas_tri_release(S.ticket);  // waits for the meshes
    S.ticket = nullptr;
    S.t_done = clk::now();
    for (int i = 0; i < n; i++) {
        int nt[2] = {0, 0};
        if (S.status[i] == ELAS_OK) {
            for (int k = 0; k < 2; k++) {
                const int r = S.jobs[S.job_of[2 * i + k]].result;
                nt[k] = r < 0 ? 0 : r;
                if (r < 0) S.status[i] = ELAS_ERR_TOO_FEW_POINTS;
            }
            if (nt[0] > S.dd[i].tricap || nt[1] > S.dd[i].tricap)
                return fail(ELAS_ERR_HIP, "triangle count exceeds capacity");
        }
        if (S.status[i] != ELAS_OK) nt[0] = nt[1] = 0;  // a pair without a mesh rasterises nothing
        S.ntri0[i] = nt[0];
        S.ntri1[i] = nt[1];
        S.pairs[i].ntri[0] = nt[0];
        S.pairs[i].ntri[1] = nt[1];
    }
    build_prefix(prm, S);
    // behind an armed gate the real table goes to its own pinned copy: a gate
    // that timed out may be copying the placeholder in hpairs[1] right now
    const int tab = S.gate_armed ? 2 : 1;
    memcpy(S.hpairs[tab], S.pairs.data(), sizeof(ElasPair) * n);
    memcpy(S.hprefix[tab], S.prefix.data(), sizeof(int32_t) * S.prefix.size());
    S.hostpre[1] = S.prefix;
    S.t_tri1 = clk::now();
    if (S.gate_armed) {
        // phase B is queued: meshes and table are in place, open the gate
        gate_release(S);
        return slot_launch_copies(S);
    }
    if (S.gated)  // the gated form of phase B has no table copy of its own
        HIPCHECK(hipMemcpyAsync(S.dpairs[1], S.hpairs[1], S.tab_pairs + sizeof(int32_t) * S.prefix.size(),
                                hipMemcpyHostToDevice, S.stream));
    return slot_launch_b(S);
}

// Launches phase B (graph or eager), then the result copies of the chunk.
int elas_ctx::slot_launch_b(ElasSlot& S) {
    hipStream_t s1 = S.stream;
    HIPCHECK(hipEventRecord(S.ev[6], s1));
    if (use_graphs) {
        hipGraphExec_t ex = nullptr;
        int r = slot_graph(S, 1, &ex);
        if (r != ELAS_OK) return r;
        HIPCHECK(hipGraphLaunch(ex, s1));
    } else {
        int r = enqueue_b(S, true);
        if (r != ELAS_OK) return r;
    }
    HIPCHECK(hipEventRecord(S.ev[10], s1));
    return slot_launch_copies(S);
}

int elas_ctx::slot_launch_copies(ElasSlot& S) {
    const int n = S.n;
    hipStream_t s1 = S.stream;
    for (int i = 0; i < n; i++)
        if (!S.specs[i].dev_out) {
            HIPCHECK(hipMemcpyAsync(S.specs[i].D1, S.pairs[i].dout[0], 4 * S.dd[i].npx, hipMemcpyDeviceToHost, s1));
            HIPCHECK(hipMemcpyAsync(S.specs[i].D2, S.pairs[i].dout[1], 4 * S.dd[i].npx, hipMemcpyDeviceToHost, s1));
        }
    HIPCHECK(hipEventRecord(S.ev[13], s1));
    S.t_blaunch = clk::now();
    S.phase = SLOT_B;
    return ELAS_OK;
}

// Phase B on the pipeline stream: table upload (unless a gate kernel copies
// the table), plane records, raster, dense matching, LR check, post filters.  Every launch depends only on the chunk's
// shapes and parameters (the plane kernel reads its triangle count from the
// table), so the sequence can be captured once per shape bucket.
int elas_ctx::enqueue_b(ElasSlot& S, bool timed) {
    const int n = S.n;
    hipStream_t s1 = S.stream;
    auto ev = [&](int i) { return timed ? hipEventRecord(S.ev[i], s1) : hipSuccess; };
    // the table goes up here, or, behind a gate, with the gate kernel
    if (!S.gated)
        HIPCHECK(hipMemcpyAsync(S.dpairs[1], S.hpairs[1], S.tab_pairs + sizeof(int32_t) * S.prefix.size(),
                                hipMemcpyHostToDevice, s1));
    const int cap256 = std::max(1, num_cu) * 8;
    auto g = [&](const ElasLaunch& L, int cap) { return std::min(L.total, cap); };
    HIPCHECK(ev(15));
    {
        // plane records, boxes and culling masks, then the tile raster that
        // writes the priors.  Triangle counts are data: the plane grid comes
        // from the largest possible count (a shape property) and the kernel
        // reads the actual count from the table
        long maxchunks = 0;
        for (int i = 0; i < n; i++) maxchunks += 2L * cdiv(S.dd[i].tricap, 64);
        elas_launch_plane(launch_of(S, 1, S_PLANE), (int)std::max(1L, std::min<long>(2 * cap256, maxchunks)),
                          S.plane_lds, s1);
        ElasLaunch L = launch_of(S, 1, S_RTILES);
        if (L.total) elas_launch_rtile(L, g(L, cap256), s1);
    }
    HIPCHECK(ev(7));
    for (int c = 1; c <= 2; c++) {
        ElasLaunch L = launch_of(S, 1, c == 1 ? S_DENSE : S_DENSE2);
        if (L.total) elas_launch_dense(L, S.kpl, c, dense_grid(L.total, num_cu), S.dense_lds[c - 1], s1);
    }
    HIPCHECK(ev(8));
    { ElasLaunch L = launch_of(S, 1, S_PIXELS); if (L.total) elas_launch_lr(L, kp, g(L, cap256), s1); }
    if (debug) { ElasLaunch L = launch_of(S, 1, S_PIXELS); if (L.total) elas_launch_mean(L, g(L, cap256), s1); }
    HIPCHECK(ev(9));
    // post-processing on D1 (and D2): speckle, gap interpolation, adaptive mean, median
    {
        const ElasLaunch Lt = launch_of(S, 1, S_POST_TILES), Lp = launch_of(S, 1, S_PIXELS);
        if (kp.speckle_size > 0 && Lt.total) elas_launch_speckle(Lt, Lp, kp, cap256, s1);
        if (kp.ipol_gap > 0 && Lt.total) {
            int mw = 1, mh = 1;
            for (int i = 0; i < n; i++) { mw = std::max(mw, S.specs[i].W); mh = std::max(mh, S.specs[i].H); }
            elas_launch_gap(launch_of(S, 1, S_POST_ROWS), launch_of(S, 1, S_POST_COLS), kp, cap256,
                            4 * (size_t)mw, 4 * (size_t)mh, s1);
        }
        int in_tmp = 0;
        if (kp.adaptive_mean && Lt.total) { elas_launch_sep(Lt, kp, false, in_tmp, cap256, s1); in_tmp ^= 1; }
        if (kp.median && Lt.total) { elas_launch_sep(Lt, kp, true, in_tmp, cap256, s1); in_tmp ^= 1; }
        if (in_tmp && Lp.total) elas_launch_post_copy(Lp, kp, cap256, s1);
    }
    HIPCHECK(hipGetLastError());
    return ELAS_OK;
}

// The executable graph of phase `which` (0: A, 1: B) for the chunk's shape
// bucket, captured from enqueue_a/enqueue_b on first use.  The key holds
// everything a launch depends on: pair shapes, where buffers live, the
// parameter generation and the debug flag; the tables and pinned buffers
// the graph reads are the slot's own, at fixed addresses.
int elas_ctx::slot_graph(ElasSlot& S, int which, hipGraphExec_t* out) {
    std::vector<int64_t> key;
    key.reserve(6 + 5 * (size_t)S.n);
    key.push_back(S.n);
    key.push_back((int64_t)param_gen);
    key.push_back((debug ? 1 : 0) | (S.gated ? 2 : 0));
    key.push_back((int64_t)(uintptr_t)S.arena);
    key.push_back((int64_t)(uintptr_t)S.table);
    for (int i = 0; i < S.n; i++) {
        key.push_back(S.specs[i].W);
        key.push_back(S.specs[i].H);
        key.push_back(S.specs[i].dmax);
        key.push_back((S.specs[i].dev_in ? 1 : 0) | (S.specs[i].dev_out ? 2 : 0));
        key.push_back(((int64_t)S.pairs[i].dsh << 32) | (S.pairs[i].dcpw * 65536 + S.pairs[i].dspc));
    }
    ElasGraphEntry* e = nullptr;
    for (auto& x : S.graphs)
        if (x.key == key) e = &x;
    if (!e) {
        if (S.graphs.size() >= ELAS_GRAPHS_PER_SLOT) {
            auto it = std::min_element(S.graphs.begin(), S.graphs.end(),
                                       [](const ElasGraphEntry& a, const ElasGraphEntry& b) { return a.use < b.use; });
            for (auto& x : it->exec)
                if (x) (void)hipGraphExecDestroy(x);
            S.graphs.erase(it);
        }
        S.graphs.emplace_back();
        e = &S.graphs.back();
        e->key = key;
    }
    e->use = ++S.graph_clock;
    if (!e->exec[which]) {
        hipStream_t s1 = S.stream;
        HIPCHECK(hipStreamBeginCapture(s1, hipStreamCaptureModeThreadLocal));
        int r = which == 0 ? enqueue_a(S, false) : enqueue_b(S, false);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(s1, &graph);
        if (r != ELAS_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return r;
        }
        if (ec != hipSuccess || !graph) return fail(ELAS_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&e->exec[which], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            e->exec[which] = nullptr;
            return fail(ELAS_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
        S.graph_captures++;
    }
    *out = e->exec[which];
    return ELAS_OK;
}

void elas_ctx::slot_drop_graphs(ElasSlot& S) {
    for (auto& x : S.graphs)
        for (auto& g : x.exec)
            if (g) (void)hipGraphExecDestroy(g);
    S.graphs.clear();
}

bool elas_ctx::slot_b_ready(ElasSlot& S) { return hipEventQuery(S.ev[13]) == hipSuccess; }

int elas_ctx::slot_finish(ElasSlot& S, int32_t* statuses) {
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(poll_event(S.ev[13]));
    S.t_end = clk::now();
    if (S.gate_used && __atomic_load_n(&S.gate[1], __ATOMIC_ACQUIRE) != 0u)
        return fail(ELAS_ERR_HIP, "phase B gate timed out waiting for the host step");
    auto el = [&](int a, int b) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, S.ev[a], S.ev[b]);
        return ms;
    };
    // per-stage times exist only for eagerly launched phases (a captured
    // graph has no events between its kernels): -1 otherwise
    const bool st = !S.graphed;
    S.times[ELAS_T_DESCRIPTOR] = st ? el(1, 2) : -1.0f;
    S.times[ELAS_T_SUPPORT] = st ? el(2, 3) : -1.0f;
    S.times[ELAS_T_FILTER] = st ? el(3, 4) : -1.0f;
    S.times[ELAS_T_COMPACT] = st ? el(4, 5) : -1.0f;
    S.times[ELAS_T_GRID] = el(11, 12);
    S.times[ELAS_T_TRIANGLE] = std::chrono::duration<float, std::milli>(S.t_tri1 - S.t_tri0).count();
    S.times[ELAS_T_PLANE] = st ? el(15, 7) : -1.0f;
    S.times[ELAS_T_DENSE] = st ? el(7, 8) : -1.0f;
    S.times[ELAS_T_LR] = st ? el(8, 9) : -1.0f;
    S.times[ELAS_T_POST] = st ? el(9, 10) : -1.0f;
    S.times[ELAS_T_TOTAL] = std::chrono::duration<float, std::milli>(S.t_end - S.t_start).count();
    S.times[ELAS_T_PHASE_A] = el(0, 5);
    S.times[ELAS_T_PHASE_B] = el(6, 10);
    auto msd = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<float, std::milli>(b - a).count(); };
    S.times[ELAS_T_SP_DOWNLOAD] = msd(S.t_tri0, S.t_sp);
    S.times[ELAS_T_TRI_WALL] = msd(S.t_sub, S.t_done);
    float tj = 0.f;
    for (const ElasTriJob& j : S.jobs) tj = std::max(tj, (float)((j.t_end - j.t_begin) * 1e3));
    S.times[ELAS_T_TRI_JOB] = tj;
    S.times[ELAS_T_HOST_A] = msd(S.t_start, S.t_alaunch);
    S.times[ELAS_T_HOST_B] = msd(S.t_done, S.t_blaunch);
    S.times[ELAS_T_GPU_SPAN] = el(0, 13);
    S.times[ELAS_T_AB_GAP] = el(5, 6);
    if (statuses)
        for (int i = 0; i < S.n; i++) statuses[i] = S.status[i];
    if (autotune && force_cpw == 0 && force_dsh == 0 && !S.external)
        for (int i = 0; i < S.n; i++)
            if (S.status[i] == ELAS_OK && !tuned.count({S.specs[i].W, S.specs[i].H, S.specs[i].dmax})) {
                const int r = tune(S, i);
                if (r != ELAS_OK) return r;
            }
    S.phase = SLOT_IDLE;
    last = &S;
    return ELAS_OK;
}

// ---- tilings and the autotuner --------------------------------------------------

// Tiling of a pair shape: the tuned one; else, for dense, two cells per wave
// where the rings fit in 64 KiB, with bands cut into chunks so that the units
// fill about four workgroups per CU, and for the descriptor the tallest strip
// that still gives every SIMD two waves.  force_cpw / force_spc / force_dsh
// (> 0) override (two cells per wave still only where they fit).
ElasTiles elas_ctx::tiles(const ElasDims& d, int W, int H, int dmax) const {
    const int gs = prm.grid_size;
    ElasTiles t;
    auto it = tuned.find({W, H, dmax});
    if (it != tuned.end()) t = it->second;
    if (it == tuned.end() || force_cpw > 0) {
        const size_t lds2 = dense_lds(gs, dmax, 2);
        t.cpw = (force_cpw == 0 || force_cpw == 2) && lds2 <= std::min<size_t>(max_lds, 65536) ? 2 : 1;
        const int segs = dense_segs(d.gw, t.cpw);
        if (force_spc > 0) {
            t.spc = std::min(force_spc, segs);
        } else {
            const long bands = (long)d.gh * cdiv(gs, dense_rows(gs));
            const long want = 4L * std::max(1, num_cu);
            const int chunks = (int)std::max(1L, std::min<long>(segs, (want + bands - 1) / bands));
            t.spc = cdiv(segs, chunks);
        }
        t.us = -1.0f;
    }
    if (force_dsh > 0) {
        t.dsh = force_dsh;
        t.desc_us = -1.0f;
    } else if (it == tuned.end()) {
        t.dsh = 6;
        for (int sh : {18, 12})
            if (desc_units(W, H, sh) >= 8 * std::max(1, num_cu)) {
                t.dsh = sh;
                break;
            }
        t.desc_us = -1.0f;
    }
    return t;
}

// Times every candidate tiling of pair i's dense launch and every descriptor
// strip height on that pair, and keeps the fastest of each for the pair's
// shape.  The inputs of the call that just finished (images, descriptors,
// priors, candidate masks) are still in place, and every tiling recomputes
// the same descriptors and disparities from them, so the results of the call
// are unchanged.  Dense candidates: 1 or 2 cells per wave (where the rings
// fit) x every distinct number of chunks per band.
int elas_ctx::tune(ElasSlot& S, int i) {
    const ElasPairSpec& sp = S.specs[i];
    const ElasDims& d = S.dd[i];
    const int gs = prm.grid_size;
    std::vector<ElasTiles> cand;
    for (int cpw = 1; cpw <= 2; cpw++) {
        // two cells per wave only within 64 KiB (one cell always runs)
        if (dense_lds(gs, sp.dmax, cpw) > (cpw == 1 ? max_lds : std::min<size_t>(max_lds, 65536))) continue;
        const int segs = dense_segs(d.gw, cpw);
        int prev = -1;
        for (int spc = 1; spc <= segs; spc++) {
            const int chunks = cdiv(segs, spc);
            if (chunks == prev) continue;  // same units as a shorter chunk
            prev = chunks;
            ElasTiles t;
            t.cpw = cpw;
            t.spc = spc;
            cand.push_back(t);
        }
    }
    // device records: the dense candidates, then the pair as it is (for the
    // descriptor); prefixes [0, units] of the dense candidates, then of the
    // descriptor strip heights
    const size_t nc = cand.size(), nrec = nc + 1, npre = nc + 3;
    const size_t pre_off = al(sizeof(ElasPair) * nrec);
    const size_t need = pre_off + al(8 * npre);
    HIPCHECK(hipSetDevice(device));
    if (S.tune_bytes < need) {
        if (S.tune) (void)hipFree(S.tune);
        S.tune = nullptr;
        S.tune_bytes = 0;
        HIPCHECK(hipMalloc(&S.tune, need));
        S.tune_bytes = need;
    }
    std::vector<char> h(need, 0);
    std::vector<int> units(npre);
    for (size_t k = 0; k < nrec; k++) {
        ElasPair P = S.pairs[i];
        if (k < nc) {
            P.dcpw = cand[k].cpw;
            P.dspc = cand[k].spc;
            units[k] = dense_units(d, gs, P.dcpw, P.dspc);
        }
        memcpy(h.data() + k * sizeof(ElasPair), &P, sizeof(P));
    }
    for (int j = 0; j < 3; j++) units[nc + j] = desc_units(sp.W, sp.H, kDescRows[j]);
    for (size_t k = 0; k < npre; k++) {
        int32_t* pf = (int32_t*)(h.data() + pre_off) + 2 * k;
        pf[0] = 0;
        pf[1] = units[k];
    }
    hipStream_t s1 = S.stream;
    HIPCHECK(hipMemcpyAsync(S.tune, h.data(), need, hipMemcpyHostToDevice, s1));
    const int reps = 3;
    // average microseconds of `reps` launches after a warm one
    auto time_us = [&](auto&& launch, float* us) -> int {
        launch();
        HIPCHECK(hipEventRecord(S.ev[16], s1));
        for (int r = 0; r < reps; r++) launch();
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(S.ev[17], s1));
        HIPCHECK(hipEventSynchronize(S.ev[17]));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, S.ev[16], S.ev[17]));
        *us = ms * 1e3f / reps;
        return ELAS_OK;
    };
    auto launch_of_k = [&](size_t rec, size_t pre) {
        ElasLaunch L;
        L.pairs = (const ElasPair*)S.tune + rec;
        L.prefix = (const int32_t*)((const char*)S.tune + pre_off) + 2 * pre;
        L.npairs = 1;
        L.total = units[pre];
        return L;
    };
    for (size_t k = 0; k < nc; k++) {
        const ElasLaunch L = launch_of_k(k, k);
        const size_t lds = dense_lds(gs, sp.dmax, cand[k].cpw);
        const int r = time_us([&] { elas_launch_dense(L, S.kpl, cand[k].cpw, dense_grid(L.total, num_cu), lds, s1); },
                              &cand[k].us);
        if (r != ELAS_OK) return r;
    }
    size_t best = 0;
    for (size_t k = 1; k < nc; k++)
        if (cand[k].us < cand[best].us) best = k;
    ElasTiles t = nc ? cand[best] : tiles(d, sp.W, sp.H, sp.dmax);
    const int cap256 = std::max(1, num_cu) * 8;
    for (int j = 0; j < 3; j++) {
        const ElasLaunch L = launch_of_k(nc, nc + j);
        float us = 0.f;
        const int r = time_us([&] { elas_launch_descriptor(L, kDescRows[j], std::min(L.total, cap256), s1); }, &us);
        if (r != ELAS_OK) return r;
        if (j == 0 || us < t.desc_us) {
            t.dsh = kDescRows[j];
            t.desc_us = us;
        }
    }
    tuned[{sp.W, sp.H, sp.dmax}] = t;
    tunings++;
    return ELAS_OK;
}

// Returns a slot to idle after an error: waits for its queued GPU work and
// its Triangle jobs (both write into the slot's buffers).
void elas_ctx::slot_abort(ElasSlot& S) {
    (void)hipSetDevice(device);
    gate_release(S);
    if (S.ticket) elas_tri_release(S.ticket);
    S.ticket = nullptr;
    if (S.stream) (void)hipStreamSynchronize(S.stream);
    if (S.stream2) (void)hipStreamSynchronize(S.stream2);
    S.phase = SLOT_IDLE;
}

// ---- drivers -----------------------------------------------------------------

int elas_ctx::run(const ElasPairSpec* sp, int n, int32_t* statuses) {
    if (slots.empty()) return fail(ELAS_ERR_INVALID_ARGUMENT, "context has no slot");
    ElasSlot& S = *slots[0];
    std::vector<int32_t> st(n > 0 ? n : 1, ELAS_OK);
    int r = slot_prepare(S, sp, n);
    S.gated = gate_b;  // one slot: phase B may wait on the stream behind its gate
    if (r == ELAS_OK) r = slot_phase_a(S);
    if (r == ELAS_OK) r = slot_start_tri(S);
    if (r == ELAS_OK) r = slot_phase_b(S);
    if (r == ELAS_OK) r = slot_finish(S, st.data());
    if (r != ELAS_OK) {
        slot_abort(S);
        return r;
    }
    int ret = ELAS_OK;
    for (int i = 0; i < n; i++) {
        if (statuses) statuses[i] = st[i];
        if (st[i] != ELAS_OK && st[i] != ELAS_ERR_TOO_FEW_POINTS) ret = st[i];
    }
    if (n == 1 && st[0] != ELAS_OK) {
        ret = st[0];
        err = "fewer than 3 support points (or a degenerate mesh): D1/D2 are all -1";
    }
    return ret;
}

// The fixed slot of the no-allocation entry point: streams, events, gate and
// pinned staging now (at context creation), the GPU's scratch address from
// k_scratch.hip.  The device table takes the scratch's first 64 KiB, the
// arena the rest.
#define ELAS_FIXED_TABLE ((size_t)64 << 10)
#define ELAS_FIXED_PIN ((size_t)8 << 20)
int elas_ctx::fixed_init() {
    HIPCHECK(hipSetDevice(device));
    void* p = nullptr;
    if (elas_scratch_setup(device, &p) != 0 || !p)
        return fail(ELAS_ERR_HIP, "scratch of the no-allocation entry point: address lookup failed");
    scratch = (char*)p;
    ElasSlot* S = new (std::nothrow) ElasSlot();
    if (!S) return fail(ELAS_ERR_OUT_OF_MEMORY, "fixed slot");
    fixed = S;
    int r = slot_init(*S);
    if (r != ELAS_OK) return r;
    S->external = true;
    if (hipHostMalloc(&S->pin, ELAS_FIXED_PIN, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        S->pin = nullptr;
        (void)hipGetLastError();
        return fail(ELAS_ERR_OUT_OF_MEMORY, "pinned staging of the no-allocation entry point");
    }
    if (hipHostGetDevicePointer((void**)&S->pin_dev, S->pin, 0) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ELAS_ERR_OUT_OF_MEMORY, "pinned staging is not mapped for the device");
    }
    S->pin_bytes = ELAS_FIXED_PIN;
    S->table = scratch;
    S->table_bytes = ELAS_FIXED_TABLE;
    S->arena = scratch + ELAS_FIXED_TABLE;
    S->arena_bytes = ELAS_SCRATCH_BYTES - ELAS_FIXED_TABLE;
    return ELAS_OK;
}

// One device-resident pair through the fixed slot.  No hipMalloc, hipFree,
// pinned allocation, graph instantiation or autotuning happens here: the
// slot's buffers were made by fixed_init, phases A and B are launched
// directly (phase B behind the gate), and a shape's tiling is the tuned one
// if an earlier call on this context tuned it, else the heuristic.
int elas_ctx::run_fixed(const ElasPairSpec& sp) {
    if (!fixed || !scratch) return fail(ELAS_ERR_HIP, "no-allocation entry point not set up: " + fixed_err);
    ElasSlot& S = *fixed;
    size_t arena_need = 0, pin_need = 0;
    int r = slot_need(&sp, 1, &arena_need, &pin_need);
    if (r != ELAS_OK) return r;
    if (arena_need > S.arena_bytes || pin_need > S.pin_bytes)
        return fail(ELAS_ERR_UNSUPPORTED, "pair needs " + std::to_string(arena_need) + " B of device scratch and " +
                                              std::to_string(pin_need) + " B of pinned staging; the no-allocation "
                                              "entry point holds " + std::to_string(S.arena_bytes) + " and " +
                                              std::to_string(S.pin_bytes));
    int32_t st = ELAS_OK;
    r = slot_prepare(S, &sp, 1);
    S.gated = true;  // direct launches only (a graph would be instantiated)
    if (r == ELAS_OK) r = slot_phase_a(S);
    if (r == ELAS_OK) r = slot_start_tri(S);
    if (r == ELAS_OK) r = slot_phase_b(S);
    if (r == ELAS_OK) r = slot_finish(S, &st);
    if (r != ELAS_OK) {
        slot_abort(S);
        return r;
    }
    if (st != ELAS_OK) err = "fewer than 3 support points (or a degenerate mesh): D1/D2 are all -1";
    return st;
}

// Chunks of the batch move through the slots: a slot whose phase A is done
// hands its support points to the Triangle workers, a slot whose meshes are
// done enqueues phase B, a finished slot takes the next chunk.  Each slot has
// its own streams, so the GPU phases of different chunks overlap each other
// and the host step.
int elas_ctx::run_batch(const ElasPairSpec* sp, int n, int32_t* statuses, int chunk) {
    if (n <= 0 || !sp) return fail(ELAS_ERR_INVALID_ARGUMENT, "empty batch");
    if (chunk <= 0 || chunk > n) chunk = n;
    const int nchunks = (n + chunk - 1) / chunk;
    if (nchunks == 1 || slots.size() < 2) {
        // one slot: chunks in sequence
        for (int c = 0; c < nchunks; c++) {
            const int c0 = c * chunk, cn = std::min(chunk, n - c0);
            int r = run(sp + c0, cn, statuses ? statuses + c0 : nullptr);
            if (r != ELAS_OK && !(cn == 1 && r == ELAS_ERR_TOO_FEW_POINTS)) return r;
        }
        return ELAS_OK;
    }
    // reserve every slot for the largest chunk before anything is in flight
    size_t a_max = 0, p_max = 0;
    for (int c = 0; c < nchunks; c++) {
        size_t a = 0, p = 0;
        const int c0 = c * chunk, cn = std::min(chunk, n - c0);
        int r = slot_need(sp + c0, cn, &a, &p);
        if (r != ELAS_OK) return r;
        a_max = std::max(a_max, a);
        p_max = std::max(p_max, p);
    }
    for (ElasSlot* S : slots) {
        int r = slot_reserve(*S, a_max, p_max, chunk);
        if (r != ELAS_OK) return r;
    }
    std::vector<int> chunk_of(slots.size(), -1);
    int next = 0, done = 0, ret = ELAS_OK;
    while (done < nchunks && ret == ELAS_OK) {
        bool progressed = false;
        for (size_t k = 0; k < slots.size() && ret == ELAS_OK; k++) {
            ElasSlot& S = *slots[k];
            int r = ELAS_OK;
            if (S.phase == SLOT_IDLE && next < nchunks) {
                const int c0 = next * chunk, cn = std::min(chunk, n - c0);
                chunk_of[k] = next++;
                r = slot_prepare(S, sp + c0, cn);
                if (r == ELAS_OK) r = slot_phase_a(S);
                progressed = true;
            } else if (S.phase == SLOT_A && slot_a_ready(S)) {
                r = slot_start_tri(S);
                progressed = true;
            } else if (S.phase == SLOT_TRI && slot_tri_ready(S)) {
                r = slot_phase_b(S);
                progressed = true;
            } else if (S.phase == SLOT_B && slot_b_ready(S)) {
                r = slot_finish(S, statuses ? statuses + (size_t)chunk_of[k] * chunk : nullptr);
                done++;
                progressed = true;
            }
            if (r != ELAS_OK) ret = r;
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    if (ret != ELAS_OK) {
        const std::string msg = err;
        for (ElasSlot* S : slots) slot_abort(*S);
        err = msg;
        return ret;
    }
    if (statuses)
        for (int i = 0; i < n; i++)
            if (statuses[i] != ELAS_OK && statuses[i] != ELAS_ERR_TOO_FEW_POINTS) return statuses[i];
    return ELAS_OK;
}

int elas_ctx::fetch(int stage, void* dst, size_t bytes, size_t* stage_bytes) {
    if (!last || last->pairs.empty()) return fail(ELAS_ERR_INVALID_ARGUMENT, "no process call yet");
    const ElasSlot& S = *last;
    const ElasPair& P = S.pairs[0];
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
        case ELAS_STAGE_SP0: src = P.sp[0]; sz = 8 * (size_t)S.nsp[0]; break;
        case ELAS_STAGE_SP1: src = P.sp[1]; sz = 8 * (size_t)S.nsp[0]; break;
        case ELAS_STAGE_GRID0: src = P.glist[0]; sz = 4 * (size_t)P.gw * P.gh * (P.dmax + 2); break;
        case ELAS_STAGE_GRID1: src = P.glist[1]; sz = 4 * (size_t)P.gw * P.gh * (P.dmax + 2); break;
        case ELAS_STAGE_TRI0:
        case ELAS_STAGE_TRI1: {
            // the meshes live in pinned memory, as 16-bit indices when P.tri16
            const int k = stage == ELAS_STAGE_TRI1;
            const size_t nt = (size_t)(k ? S.ntri1[0] : S.ntri0[0]);
            const int32_t* h = S.htri + S.tri_off[0] + (size_t)k * 3 * S.dd[0].tricap;
            if (stage_bytes) *stage_bytes = 12 * nt;
            if (!dst || bytes == 0 || nt == 0) return ELAS_OK;
            const size_t ne = std::min(bytes / 4, 3 * nt);
            int32_t* o = (int32_t*)dst;
            for (size_t e = 0; e < ne; e++) o[e] = P.tri16 ? (int32_t)((const uint16_t*)h)[e] : h[e];
            return ELAS_OK;
        }
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
    HIPCHECK(hipMemcpy(dst, src, std::min(bytes, sz), hipMemcpyDefault));  // the meshes live in pinned memory
    return ELAS_OK;
}

int elas_ctx::time_kernel(int kernel, int reps, float* avg_ms) {
    if (!last || last->pairs.empty()) return fail(ELAS_ERR_INVALID_ARGUMENT, "no process call yet");
    if (reps < 1 || reps > 100000 || !avg_ms) return fail(ELAS_ERR_INVALID_ARGUMENT, "reps must be in [1, 100000]");
    ElasSlot& S = *last;
    HIPCHECK(hipSetDevice(device));
    const int cap256 = std::max(1, num_cu) * 8;
    auto g = [&](const ElasLaunch& L, int cap) { return std::min(L.total, cap); };
    ElasLaunch L;
    int dense_cpw = 1, desc_sh = 18;
    switch (kernel) {
        case ELAS_K_DESCRIPTOR:
            // the launch of the first strip height any pair uses
            for (int sh : {18, 12, 6}) {
                L = launch_of(S, 0, sh == 18 ? S_DESC : sh == 12 ? S_DESC12 : S_DESC6);
                desc_sh = sh;
                if (L.total) break;
            }
            break;
        case ELAS_K_SUPPORT: L = launch_of(S, 0, S_SUPPORT); break;
        case ELAS_K_DENSE:
            // the launch of the pairs tiled with one cell per wave, else two
            L = launch_of(S, 1, S_DENSE);
            if (L.total == 0) {
                L = launch_of(S, 1, S_DENSE2);
                dense_cpw = 2;
            }
            break;
        case ELAS_K_LR: L = launch_of(S, 1, S_PIXELS); break;
        default: return fail(ELAS_ERR_INVALID_ARGUMENT, "unknown kernel");
    }
    if (L.total == 0) return fail(ELAS_ERR_INVALID_ARGUMENT, "kernel has no work in the last call");
    hipStream_t s1 = S.stream;
    auto launch = [&]() {
        switch (kernel) {
            case ELAS_K_DESCRIPTOR: elas_launch_descriptor(L, desc_sh, g(L, cap256), s1); break;
            case ELAS_K_SUPPORT:
                elas_launch_support(L, kp, g(L, cap256), support_lds(prm.candidate_stepsize, S.maxd, true), s1);
                break;
            case ELAS_K_DENSE:
                elas_launch_dense(L, S.kpl, dense_cpw, dense_grid(L.total, num_cu), S.dense_lds[dense_cpw - 1], s1);
                break;
            case ELAS_K_LR: elas_launch_lr(L, kp, g(L, cap256), s1); break;
        }
    };
    launch();  // warm
    HIPCHECK(hipEventRecord(S.ev[16], s1));
    for (int i = 0; i < reps; i++) launch();
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(S.ev[17], s1));
    HIPCHECK(hipEventSynchronize(S.ev[17]));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, S.ev[16], S.ev[17]));
    *avg_ms = ms / reps;
    return ELAS_OK;
}
