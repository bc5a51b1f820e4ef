// elas_engine.h -- host engine behind the C ABI.
//
// A context owns one device and a set of pipeline slots.  A slot carries one
// chunk of pairs through the three phases of the path:
//   A  (GPU)  descriptors, support matching, filters, compaction, grid;
//   T  (host) Triangle on the worker pool (non-blocking);
//   B  (GPU)  plane prior, dense matching, left/right check;
// with its own streams, HBM arena, device tables and pinned staging, so that
// several chunks can be in different phases at once.  A plain call uses
// slot 0 start to finish; a batch or the serving loop drives several slots.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <array>
#include <chrono>
#include <map>
#include <mutex>
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
    int64_t tag;  // serving request id (-1 otherwise)
};

// Derived sizes of one pair.
struct ElasDims {
    int32_t cw, ch, gw, gh, words, nblk, tricap;
    int32_t sshift, nsx, nst, scap;  // raster culling super-tiles and their list capacity
    size_t npx, ncan, ncell, nspcap;  // nspcap: support point capacity (lattice + corners)
};

enum ElasSlotPhase { SLOT_IDLE = 0, SLOT_A = 1, SLOT_TRI = 2, SLOT_B = 3 };

// Captured phases of one shape bucket (key: pair shapes, buffer placement,
// parameter generation, debug flag, arena/table addresses).
#define ELAS_GRAPHS_PER_SLOT 8
struct ElasGraphEntry {
    std::vector<int64_t> key;
    hipGraphExec_t exec[2] = {nullptr, nullptr};  // phase A, phase B
    uint64_t use = 0;
};

struct ElasSlot {
    hipStream_t stream = nullptr;   // phases A and B
    hipStream_t stream2 = nullptr;  // disparity grid, overlapped with the host step
    hipEvent_t ev[18] = {};
    // device memory
    void* arena = nullptr;
    size_t arena_bytes = 0;
    void* table = nullptr;  // 2 x (ElasPair[n] + prefix[S_NUM_STAGES][n+1] + nsp[n])
    size_t table_bytes = 0;
    // pinned host staging
    void* pin = nullptr;
    char* pin_dev = nullptr;  // device address of pin (mapped, coherent: kernels write results there)
    size_t pin_bytes = 0;
    void* tune = nullptr;  // dense tiling candidates: pair records + prefixes
    size_t tune_bytes = 0;

    // the chunk in flight
    int phase = SLOT_IDLE;
    int n = 0;
    std::vector<ElasPairSpec> specs;
    std::vector<ElasDims> dd;
    std::vector<size_t> base, sp_off, tri_off;
    std::vector<ElasPair> pairs;
    std::vector<int32_t> prefix;
    std::vector<int32_t> nsp, ntri0, ntri1, status;
    int32_t last_nsp = 0;  // support points of the slot's last single-pair call
    int maxd = 0;
    size_t dense_lds[2] = {0, 0};  // dynamic LDS of the dense launches (1, 2 cells per wave)
    size_t plane_lds = 0;          // dynamic LDS of the plane launch (super-tile masks)
    ElasKParams kpl;
    size_t tab_pairs = 0, tab_prefix = 0, tab_one = 0, pin_sp = 0, pin_tri = 0;
    ElasPair* dpairs[2] = {nullptr, nullptr};
    int32_t* dprefix[2] = {nullptr, nullptr};
    int32_t* dnsp = nullptr;
    // pinned mirrors of the tables: [0] phase A, [1] phase B, [2] phase B
    // behind a gate (the real table; [1] then holds the placeholder that a
    // timed-out gate copies instead, never written while the gate waits)
    ElasPair* hpairs[3] = {nullptr, nullptr, nullptr};
    int32_t* hprefix[3] = {nullptr, nullptr, nullptr};
    std::vector<int32_t> hostpre[2];  // host copies of the stage prefixes of tables 0 and 1
    int32_t* hnsp = nullptr;
    float* hsp = nullptr;
    int32_t* htri = nullptr;
    std::vector<ElasTriJob> jobs;
    std::vector<int> job_of;
    ElasTriTicket* ticket = nullptr;
    std::chrono::steady_clock::time_point t_start, t_alaunch, t_tri0, t_sp, t_sub, t_done, t_tri1, t_blaunch, t_end;
    float times[ELAS_T_COUNT] = {};
    // graphs of this slot's shape buckets
    std::vector<ElasGraphEntry> graphs;
    uint64_t graph_clock = 0, graph_captures = 0;
    bool graphed = false;  // the chunk in flight runs its phases as graphs
    // Phase B behind a gate (single-pair path): it is queued while the host
    // triangulates, behind a kernel that polls gate[0] and then copies the
    // phase B table from pinned memory; the host sets gate[0] once the meshes
    // and the table are written.  gate[1] reports a poll that timed out.
    uint32_t* gate = nullptr;      // mapped coherent pinned words
    uint32_t* gate_dev = nullptr;  // their device address
    bool gated = false;            // this chunk's phase B goes through the gate
    std::vector<char> tab0_up;     // bytes last copied into the phase A device table
    bool gate_armed = false;       // queued behind the gate, flag not yet set
    bool gate_used = false;        // this call's phase B went through the gate (gate[1] is its status)
    // arena and device table lie in the per-GPU scratch of the no-allocation
    // entry point: never grown, never freed, no autotuning (which allocates)
    bool external = false;
};

// Tiling of one pair shape: dense grid cells per wave per step and steps per
// chunk, descriptor strip rows, and the measured microseconds of a launch of
// each (tuned shapes; -1 otherwise).
struct ElasTiles {
    int cpw = 1, spc = 1, dsh = ELAS_DESC_SH;
    float us = -1.0f, desc_us = -1.0f;
};

struct ElasServer;

struct elas_ctx {
    int device = 0;
    int num_cu = 0;
    elas_params prm;
    ElasKParams kp;
    bool debug = false;
    bool tiny_lists = false;  // raster lists of one entry (tests the overflow path)
    int tri_workers = ELAS_TRI_INSTANCES;
    int chunk = 0;  // pairs per pipeline chunk of a batch (0: whole batch)
    bool use_graphs = true;   // phases A and B replayed as graphs per shape bucket
    bool gate_b = true;       // single-pair calls queue phase B behind a gate during Triangle
    uint64_t gate_ticks = 3000000000ull;  // the gate's poll bound (100 MHz ticks: 30 s)
    uint64_t param_gen = 0;   // bumped when parameters or the debug flag change
    std::string err;
    std::vector<ElasSlot*> slots;
    ElasSlot* last = nullptr;  // slot of the last finished call (fetch, shape info, times)
    ElasServer* server = nullptr;
    // the no-allocation entry point: its own slot (streams, events, gate and
    // pinned staging made at context creation) over the GPU's scratch
    ElasSlot* fixed = nullptr;
    char* scratch = nullptr;
    std::string fixed_err;  // why fixed_init failed, if it did
    size_t max_lds = 65536;   // dynamic LDS a workgroup may allocate on this device
    // Tilings per (W, H, disp_max): with autotune on, the first call with a
    // new shape times every candidate dense tiling and descriptor strip
    // height on that pair after the call and keeps the fastest; untuned
    // shapes use the heuristics of tiles(); force_cpw > 0 / force_dsh > 0
    // override both.
    bool autotune = true;
    int force_cpw = 0, force_spc = 0, force_dsh = 0;
    std::map<std::array<int, 3>, ElasTiles> tuned;
    uint64_t tunings = 0;
    ElasTiles tiles(const ElasDims& d, int W, int H, int dmax) const;
    int tune(ElasSlot& S, int i);

    // blocking single-chunk run on slot 0 (Elas::process and the device form)
    int run(const ElasPairSpec* specs, int n, int32_t* statuses);
    // the same for one device-resident pair in the fixed slot over the
    // scratch, with no device or pinned allocation
    int run_fixed(const ElasPairSpec& spec);
    int fixed_init();
    // pipelined run of n pairs in chunks of up to `chunk` pairs over the slots
    int run_batch(const ElasPairSpec* specs, int n, int32_t* statuses, int chunk);
    int fetch(int stage, void* dst, size_t bytes, size_t* stage_bytes);
    int time_kernel(int kernel, int reps, float* avg_ms);
    std::mutex err_m;  // the serving thread reports errors too
    int fail(int code, const std::string& msg) {
        std::lock_guard<std::mutex> g(err_m);
        err = msg;
        return code;
    }

    // phases of one slot
    int slot_init(ElasSlot& S);
    void slot_free(ElasSlot& S);
    int slot_need(const ElasPairSpec* sp, int n, size_t* arena_need, size_t* pin_need);
    int slot_reserve(ElasSlot& S, size_t arena_need, size_t pin_need, int n);
    int slot_prepare(ElasSlot& S, const ElasPairSpec* sp, int n);
    int slot_phase_a(ElasSlot& S);
    bool slot_a_ready(ElasSlot& S);
    int slot_start_tri(ElasSlot& S);
    int slot_prelaunch_b(ElasSlot& S);
    int slot_launch_b(ElasSlot& S);
    int slot_launch_copies(ElasSlot& S);
    void gate_release(ElasSlot& S);
    bool slot_tri_ready(ElasSlot& S);
    int slot_phase_b(ElasSlot& S);
    bool slot_b_ready(ElasSlot& S);
    int slot_finish(ElasSlot& S, int32_t* statuses);
    void slot_abort(ElasSlot& S);
    int enqueue_a(ElasSlot& S, bool timed);
    int enqueue_b(ElasSlot& S, bool timed);
    int slot_graph(ElasSlot& S, int which, hipGraphExec_t* out);
    void slot_drop_graphs(ElasSlot& S);
};

// Serving loop (elas_serve.cpp).
int elas_srv_start(elas_ctx* c, uint64_t budget, int max_chunk, int slots);
int elas_srv_submit(elas_ctx* c, const ElasPairSpec& s, int64_t* ticket);
int elas_srv_poll(elas_ctx* c, int64_t ticket, int32_t* status, float* latency_ms);
int elas_srv_stop(elas_ctx* c);
int elas_srv_stats(elas_ctx* c, elas_serve_stats* st);

// Derived sizes of a pair and its device footprint in a slot arena.
ElasDims elas_dims(const elas_params& p, int W, int H, int dmax);
size_t elas_pair_device_bytes(const elas_params& p, const ElasPairSpec& s, bool debug);

// Parameter checks and derived kernel parameters (prior table, key bias).
int elas_make_kparams(const elas_params* p, ElasKParams* kp, std::string* err);
