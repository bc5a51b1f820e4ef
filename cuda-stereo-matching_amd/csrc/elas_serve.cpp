// elas_serve.cpp -- continuous-batching request loop over the pipeline slots.
//
// The reference has no serving path: Elas::process (src/Elas.cu:19-120) is
// one blocking call per pair.  Here a scheduler thread owns the context's
// slots while serving:
//   * requests queue FIFO; an idle slot takes the longest prefix of the
//     queue whose device footprint fits the slot's share of the HBM budget
//     (and at most max_chunk pairs), so pairs of different sizes and
//     disparity ranges share one launch per stage;
//   * a slot that finished its GPU phase A hands its support points to the
//     Triangle workers and the scheduler moves on; requests that arrive while
//     other chunks are in flight are admitted into the next idle slot, so
//     the GPU stages of several chunks and the host step of another overlap;
//   * slot arenas are sized once to budget / slots at start and never grow,
//     which bounds the device memory of the whole loop.
// Each pair of a chunk is computed exactly as alone: the kernels index every
// buffer through the pair's own table entry and no stage mixes pairs.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "elas_engine.h"

namespace {
using clk = std::chrono::steady_clock;

struct ElasRequest {
    int64_t id = 0;
    ElasPairSpec spec{};
    size_t bytes = 0;
    int32_t status = ELAS_OK;
    bool done = false;
    clk::time_point t_submit, t_done;
};
}  // namespace

struct ElasServer {
    elas_ctx* ctx = nullptr;
    uint64_t budget = 0, cap = 0;
    int max_chunk = 32;
    std::thread th;
    std::mutex m;
    std::condition_variable cv;
    std::deque<ElasRequest*> queue;
    std::unordered_map<int64_t, ElasRequest*> reqs;
    std::vector<std::vector<ElasRequest*>> in_slot;
    std::vector<uint64_t> slot_bytes;
    int64_t next_id = 1;
    bool stop = false;
    elas_serve_stats stats{};

    void loop();
    void complete(size_t k, const int32_t* statuses, int fail_code);
};

// Marks the requests of slot k done (statuses per pair, or fail_code for all).
void ElasServer::complete(size_t k, const int32_t* statuses, int fail_code) {
    const auto now = clk::now();
    std::lock_guard<std::mutex> g(m);
    auto& v = in_slot[k];
    for (size_t i = 0; i < v.size(); i++) {
        v[i]->status = statuses ? statuses[i] : fail_code;
        v[i]->done = true;
        v[i]->t_done = now;
    }
    stats.completed += (int64_t)v.size();
    v.clear();
    slot_bytes[k] = 0;
}

void ElasServer::loop() {
    (void)hipSetDevice(ctx->device);
    std::vector<int32_t> st;
    std::vector<ElasPairSpec> specs;
    for (;;) {
        bool progressed = false, busy = false;
        for (size_t k = 0; k < ctx->slots.size(); k++) {
            ElasSlot& S = *ctx->slots[k];
            int r = ELAS_OK;
            if (S.phase == SLOT_A && ctx->slot_a_ready(S)) {
                r = ctx->slot_start_tri(S);
                progressed = true;
            } else if (S.phase == SLOT_TRI && ctx->slot_tri_ready(S)) {
                r = ctx->slot_phase_b(S);
                progressed = true;
            } else if (S.phase == SLOT_B && ctx->slot_b_ready(S)) {
                st.assign(S.n, ELAS_OK);
                r = ctx->slot_finish(S, st.data());
                if (r == ELAS_OK) complete(k, st.data(), 0);
                progressed = true;
            } else if (S.phase == SLOT_IDLE) {
                // admit the longest FIFO prefix that fits this slot
                specs.clear();
                {
                    std::lock_guard<std::mutex> g(m);
                    uint64_t bytes = 0;
                    while (!queue.empty() && (int)specs.size() < max_chunk && bytes + queue.front()->bytes <= cap) {
                        ElasRequest* q = queue.front();
                        queue.pop_front();
                        bytes += q->bytes;
                        in_slot[k].push_back(q);
                        specs.push_back(q->spec);
                    }
                    if (!specs.empty()) {
                        slot_bytes[k] = bytes;
                        stats.chunks++;
                        stats.max_chunk_pairs = std::max<int64_t>(stats.max_chunk_pairs, (int64_t)specs.size());
                        uint64_t fl = 0;
                        int64_t np = 0;
                        for (size_t j = 0; j < slot_bytes.size(); j++) {
                            fl += slot_bytes[j];
                            np += (int64_t)in_slot[j].size();
                        }
                        stats.peak_in_flight_bytes = std::max(stats.peak_in_flight_bytes, (uint64_t)fl);
                        stats.max_in_flight_pairs = std::max(stats.max_in_flight_pairs, np);
                    }
                }
                if (!specs.empty()) {
                    r = ctx->slot_prepare(S, specs.data(), (int)specs.size());
                    if (r == ELAS_OK) r = ctx->slot_phase_a(S);
                    progressed = true;
                }
            }
            if (r != ELAS_OK) {
                ctx->slot_abort(S);
                complete(k, nullptr, r);
            }
            busy = busy || S.phase != SLOT_IDLE;
        }
        if (progressed) continue;
        std::unique_lock<std::mutex> g(m);
        if (!busy && queue.empty()) {
            if (stop) return;
            cv.wait(g, [&] { return stop || !queue.empty(); });
            continue;
        }
        g.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(busy ? 20 : 0));
    }
}

int elas_srv_start(elas_ctx* c, uint64_t budget, int max_chunk, int slots) {
    if (c->server) return c->fail(ELAS_ERR_BUSY, "already serving");
    if (slots > 0 && slots != (int)c->slots.size()) {
        const int r = elas_hip_set_pipeline(c, slots, c->chunk);
        if (r != ELAS_OK) return r;
    }
    if (budget == 0) budget = (uint64_t)8 << 30;
    if (max_chunk <= 0) max_chunk = 32;
    ElasServer* sv = new ElasServer();
    sv->ctx = c;
    sv->budget = budget;
    sv->cap = budget / c->slots.size();
    sv->max_chunk = max_chunk;
    sv->in_slot.assign(c->slots.size(), {});
    sv->slot_bytes.assign(c->slots.size(), 0);
    sv->stats.slot_capacity_bytes = sv->cap;
    // size every arena once: admission keeps each chunk within it
    for (ElasSlot* S : c->slots) {
        if (S->arena_bytes < sv->cap) {
            if (S->arena) (void)hipFree(S->arena);
            S->arena = nullptr;
            S->arena_bytes = 0;
            if (hipSetDevice(c->device) != hipSuccess || hipMalloc(&S->arena, sv->cap) != hipSuccess) {
                S->arena = nullptr;
                (void)hipGetLastError();
                delete sv;
                return c->fail(ELAS_ERR_OUT_OF_MEMORY, "serving arena of " + std::to_string(sv->cap) + " bytes failed");
            }
            S->arena_bytes = sv->cap;
        }
    }
    c->server = sv;
    sv->th = std::thread([sv] { sv->loop(); });
    return ELAS_OK;
}

int elas_srv_submit(elas_ctx* c, const ElasPairSpec& s, int64_t* ticket) {
    ElasServer* sv = c->server;
    if (!sv) return c->fail(ELAS_ERR_INVALID_ARGUMENT, "not serving (call elas_hip_serve_start)");
    if (!ticket) return c->fail(ELAS_ERR_INVALID_ARGUMENT, "null ticket");
    size_t an = 0, pn = 0;
    const int r = c->slot_need(&s, 1, &an, &pn);  // validates the pair
    if (r != ELAS_OK) return r;
    if (an > sv->cap)
        return c->fail(ELAS_ERR_OUT_OF_MEMORY, "pair needs " + std::to_string(an) + " device bytes, a slot holds " +
                                                   std::to_string(sv->cap));
    ElasRequest* q = new ElasRequest();
    q->spec = s;
    q->bytes = an;
    q->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> g(sv->m);
        q->id = sv->next_id++;
        q->spec.tag = q->id;
        sv->reqs[q->id] = q;
        sv->queue.push_back(q);
        sv->stats.submitted++;
        *ticket = q->id;
    }
    sv->cv.notify_one();
    return ELAS_OK;
}

int elas_srv_poll(elas_ctx* c, int64_t ticket, int32_t* status, float* latency_ms) {
    ElasServer* sv = c->server;
    if (!sv) return c->fail(ELAS_ERR_INVALID_ARGUMENT, "not serving");
    std::lock_guard<std::mutex> g(sv->m);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return ELAS_ERR_INVALID_ARGUMENT;
    ElasRequest* q = it->second;
    if (!q->done) return ELAS_ERR_BUSY;
    if (status) *status = q->status;
    if (latency_ms) *latency_ms = std::chrono::duration<float, std::milli>(q->t_done - q->t_submit).count();
    sv->reqs.erase(it);
    delete q;
    return ELAS_OK;
}

int elas_srv_stop(elas_ctx* c) {
    ElasServer* sv = c->server;
    if (!sv) return ELAS_OK;
    {
        std::lock_guard<std::mutex> g(sv->m);
        sv->stop = true;
    }
    sv->cv.notify_all();
    sv->th.join();
    for (auto& kv : sv->reqs) delete kv.second;
    c->server = nullptr;
    delete sv;
    return ELAS_OK;
}

int elas_srv_stats(elas_ctx* c, elas_serve_stats* st) {
    ElasServer* sv = c->server;
    if (!sv || !st) return ELAS_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(sv->m);
    *st = sv->stats;
    return ELAS_OK;
}
