// Serving loop: resident references, continuously batched query sets.
//
// Threads and streams:
//   caller threads  cms_server_submit validates and copies a query set and
//                   queues it; cms_server_wait blocks on its completion;
//   worker thread   admits queued sets (FIFO, query slots + arena memory),
//                   embeds them on the embed stream, plans the band cycle and
//                   launches one step at a time on the compute stream (at
//                   most kMaxInFlight steps queued, so joining sets wait for
//                   at most that many steps);
//   delivery thread copies a finished set's output block to the caller's
//                   buffer on the copy stream and returns its memory.
// Embedding, steps and deliveries of different sets therefore overlap on the
// device. Host-side policy lives in serve_sched.cpp (no HIP).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cms_ctx.h"
#include "serve_sched.h"

using namespace cms;

namespace {

constexpr int kRing = 4;         // per-step host/device tables
constexpr int kMaxInFlight = 2;  // steps queued on the compute stream
constexpr int64_t kDefaultBudget = (int64_t)8 << 30;

int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// Byte offsets of a query set's embedding scratch inside its arena block.
struct ScratchLayout {
    int64_t rowptr, col, val, slots, ab, xroot, Ys, bm, colsum, total;
    int64_t qpad;
};

ScratchLayout scratch_layout(int nq, int64_t nnz, int64_t nlp, int64_t nseg, int64_t nrows, int64_t nwords) {
    ScratchLayout L{};
    L.qpad = ((int64_t)nq + 63) / 64 * 64;
    const int64_t nb8 = ((int64_t)nq + 7) / 8;
    int64_t at = 0;
    L.rowptr = at;
    at += align256(8 * ((int64_t)nq + 1));
    L.col = at;
    at += align256(4 * std::max<int64_t>(nnz, 1));
    L.val = at;
    at += align256(4 * std::max<int64_t>(nnz, 1));
    L.slots = at;
    at += align256(4 * (int64_t)nq);
    L.ab = at;
    at += align256(4 * nb8 * 8 * std::max<int64_t>(nlp, 1));
    L.xroot = at;
    at += align256(8 * std::max<int64_t>(nseg, 1) * L.qpad);
    L.Ys = at;
    at += align256(4 * std::max<int64_t>(nrows, 1) * L.qpad);
    L.bm = at;
    at += align256(4 * (L.qpad / kSmBlock) * std::max<int64_t>(nwords, 1));
    L.colsum = at;
    at += align256(8 * L.qpad);
    L.total = at;
    return L;
}

struct ReqHost {
    int nq = 0;
    int64_t nnz = 0;
    std::vector<int64_t> rowptr;
    std::vector<int32_t> col;
    std::vector<float> val;
    std::vector<int32_t> slots;  // query q -> slot (kept until finished: source of an async copy)
    float *out = nullptr;
    int rc = CMS_OK;
    std::string err;
    hipEvent_t ev_embedded = nullptr, ev_packed = nullptr, ev_last = nullptr;
    double t_submit = 0, t_admit = -1, t_join = -1, t_computed = -1, t_done = -1;
    int64_t join_step = -1;
    bool finished = false;
};

}  // namespace

struct cms_server {
    cms_ctx *ctx = nullptr;
    int device = 0;
    cms_server_opts opts{};
    std::unique_ptr<Scheduler> sched;

    // resident references
    CrossData cross{};
    const EmbedNode *embnodes = nullptr, *toprecs = nullptr;
    int metric = CMS_METRIC_WEIGHTED;
    int64_t nrows = 0, nwords = 0, nlp = 0, nseg = 0, nrb = 0, band_blocks = 0;
    int n_ref = 0, scale_bits = 0, variant = 1, grid_waves = 0;
    double scale = 1.0, mass_limit = 0.0;

    // device
    uint8_t *arena = nullptr;
    uint32_t *QY = nullptr, *qbm = nullptr;
    // split cross tiles: the query blocks' light bitmaps, prefixes and records
    uint32_t *qlbm = nullptr, *qpl = nullptr;
    uint4 *qcrl = nullptr;
    unsigned long long *qcolsum = nullptr, *work = nullptr, *counter = nullptr;
    uint32_t *part_acc = nullptr;  // split-tile partial sums, part_tiles x 1024 (kept zeroed)
    int64_t part_tiles = 0;
    int64_t qcrs = 0;  // record slots per query block (nwords * 32)
    float **d_slot_out[kRing] = {};
    int32_t *d_blocks[kRing] = {}, *d_dirty[kRing] = {};
    uint32_t *d_masks[kRing] = {};
    float **h_slot_out[kRing] = {};
    int32_t *h_blocks[kRing] = {}, *h_dirty[kRing] = {};
    uint32_t *h_masks[kRing] = {};
    hipEvent_t ring_begin[kRing] = {}, ring_end[kRing] = {};
    hipStream_t s_embed = nullptr, s_comp = nullptr, s_copy = nullptr;
    // delivery to pageable caller memory: two pinned staging buffers, filled
    // by DMA on s_copy and drained by a parallel host copy, alternately
    static constexpr size_t kStageBytes = (size_t)64 << 20;
    uint8_t *stage[2] = {};
    hipEvent_t ev_stage[2] = {};

    // threads
    std::thread worker, deliverer;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_deliver;
    bool stop = false, stop_delivery = false;
    int64_t next_id = 1;
    std::unordered_map<int64_t, ReqHost> reqs;
    std::vector<int64_t> embedding_ids, packing_ids;
    std::deque<int64_t> deliver_q;
    std::deque<int> inflight;  // ring entries of queued steps, in launch order
    std::deque<std::vector<int64_t>> inflight_finishers;  // per queued step: sets whose last band it is
    std::string fatal;         // device error that stopped the loop

    // statistics
    int64_t submitted = 0, completed = 0, nfailed = 0, steps = 0, arena_peak = 0;
    int slots_peak = 0;
    double kernel_ms = 0.0, pairs = 0.0;
    std::chrono::steady_clock::time_point t0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

int sfail(int code, const std::string &msg) {
    set_global_error(msg);
    return code;
}

std::string hip_msg(const char *what, hipError_t e) { return std::string(what) + ": " + hipGetErrorString(e); }

#define SRV_HIP(expr)                                        \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_msg(#expr, e_);     \
    } while (0)

// Finish a request (caller holds the lock).
void finish_locked(cms_server *s, int64_t id, int rc, const std::string &err) {
    ReqHost &r = s->reqs[id];
    r.rc = rc;
    r.err = err;
    r.finished = true;
    r.t_done = s->now();
    if (rc == CMS_OK) {
        s->completed++;
        s->pairs += (double)r.nq * s->n_ref;
    } else {
        s->nfailed++;
    }
    r.rowptr = {};
    r.col = {};
    r.val = {};
    r.slots = {};
    s->cv_done.notify_all();
}

// Embedding of an admitted set on the embed stream (no lock held; the
// request's host arrays and scheduler record are not modified meanwhile).
std::string launch_embedding(cms_server *s, ReqHost &r, int64_t scratch_off) {
    const ScratchLayout L = scratch_layout(r.nq, r.nnz, s->nlp, s->nseg, s->nrows, s->nwords);
    uint8_t *base = s->arena + scratch_off;
    int64_t *rowptr = (int64_t *)(base + L.rowptr);
    int32_t *col = (int32_t *)(base + L.col);
    float *val = (float *)(base + L.val);
    int32_t *dslots = (int32_t *)(base + L.slots);
    float *ab = (float *)(base + L.ab);
    cms_ctx *c = s->ctx;
    hipStream_t st = s->s_embed;
    SRV_HIP(hipMemcpyAsync(rowptr, r.rowptr.data(), sizeof(int64_t) * (r.nq + 1), hipMemcpyHostToDevice, st));
    if (r.nnz > 0) {
        SRV_HIP(hipMemcpyAsync(col, r.col.data(), sizeof(int32_t) * r.nnz, hipMemcpyHostToDevice, st));
        SRV_HIP(hipMemcpyAsync(val, r.val.data(), sizeof(float) * r.nnz, hipMemcpyHostToDevice, st));
    }
    SRV_HIP(hipMemcpyAsync(dslots, r.slots.data(), sizeof(int32_t) * r.nq, hipMemcpyHostToDevice, st));
    SRV_HIP(hipMemsetAsync(ab, 0, (size_t)(L.xroot - L.ab), st));
    SRV_HIP(hipMemsetAsync(base + L.bm, 0, (size_t)(L.colsum - L.bm), st));
    SRV_HIP(hipMemsetAsync(base + L.colsum, 0, (size_t)(L.total - L.colsum), st));
    SRV_HIP(launch_walk_scatter(rowptr, col, val, c->d_leaf_lp_ptr, c->d_leaf_lp, s->nlp, CMS_STORE_F32, nullptr,
                                nullptr, 0, r.nq, ab, st));
    EmbedData e{};
    e.nodes = s->embnodes;
    e.order_n = c->tree.order_n;
    e.nslots = std::max(c->tree.nslots, 1);
    e.seg_first = c->d_seg_first;
    e.seg_last = c->d_seg_last;
    e.nseg = (int)s->nseg;
    e.top = s->toprecs;
    e.ntop = (int)c->tree.top_nodes.size();
    e.ab = ab;
    e.nlp = s->nlp;
    e.n = r.nq;
    e.npad = L.qpad;
    e.nwords = s->nwords;
    e.ystride = L.qpad;
    e.scale = s->scale;
    e.Y = (uint32_t *)(base + L.Ys);
    e.bm = (uint32_t *)(base + L.bm);
    e.xroot = (double *)(base + L.xroot);
    e.unweighted = s->metric == CMS_METRIC_UNWEIGHTED ? 1 : 0;
    e.sums_only = 0;
    e.colsum = (unsigned long long *)(base + L.colsum);
    e.colsum_f64 = nullptr;
    SRV_HIP(launch_embed(e, st));
    SRV_HIP(hipEventRecord(r.ev_embedded, st));
    return std::string();
}

struct StepLaunch {
    SchedStep plan;
    int ring = 0;
    int nqa = 0, ndirty = 0;
    struct Join {
        int64_t id;
        ReqHost *r;
        int64_t scratch_off;
    };
    std::vector<Join> joins;
    std::vector<ReqHost *> finishers;
};

// Queue one step on the compute stream (no lock held).
std::string launch_step(cms_server *s, StepLaunch &L) {
    hipStream_t st = s->s_comp;
    const int k = L.ring;
    for (auto &j : L.joins) {
        const ScratchLayout SL = scratch_layout(j.r->nq, j.r->nnz, s->nlp, s->nseg, s->nrows, s->nwords);
        uint8_t *base = s->arena + j.scratch_off;
        SRV_HIP(hipStreamWaitEvent(st, j.r->ev_embedded, 0));
        SRV_HIP(launch_pack((const uint32_t *)(base + SL.Ys), SL.qpad, j.r->nq, (const int32_t *)(base + SL.slots),
                            s->nrows, s->QY, s->opts.slots, (const unsigned long long *)(base + SL.colsum), s->qcolsum,
                            st));
        SRV_HIP(hipEventRecord(j.r->ev_packed, st));
    }
    if (L.ndirty > 0) {
        SRV_HIP(hipMemcpyAsync(s->d_dirty[k], s->h_dirty[k], sizeof(int32_t) * L.ndirty, hipMemcpyHostToDevice, st));
        SRV_HIP(hipMemcpyAsync(s->d_masks[k], s->h_masks[k], sizeof(uint32_t) * L.ndirty, hipMemcpyHostToDevice, st));
        SRV_HIP(launch_qbm_build(s->QY, s->opts.slots, s->nrows, s->nwords, s->d_dirty[k], s->d_masks[k], L.ndirty,
                                 s->qbm, s->qlbm, st));
        if (s->qlbm)
            SRV_HIP(launch_light_records_listed(s->QY, s->opts.slots, s->d_dirty[k], s->d_masks[k], L.ndirty, s->qlbm,
                                                s->nwords, s->qpl, s->qcrl, s->qcrs, st));
    }
    SRV_HIP(hipMemcpyAsync(s->d_slot_out[k], s->h_slot_out[k], sizeof(float *) * s->opts.slots, hipMemcpyHostToDevice,
                           st));
    if (L.nqa > 0)
        SRV_HIP(hipMemcpyAsync(s->d_blocks[k], s->h_blocks[k], sizeof(int32_t) * L.nqa, hipMemcpyHostToDevice, st));
    const int64_t rb0 = (int64_t)L.plan.band * s->band_blocks;
    const int64_t nrb = std::max<int64_t>(0, std::min<int64_t>(s->band_blocks, s->nrb - rb0));
    SRV_HIP(hipEventRecord(s->ring_begin[k], st));
    SRV_HIP(launch_zero_counter(s->work, st));
    SRV_HIP(launch_sm_cross(s->cross, s->variant, s->d_blocks[k], L.nqa, rb0, nrb, s->d_slot_out[k], s->part_acc,
                            s->part_tiles, s->work, s->grid_waves, st));
    SRV_HIP(hipEventRecord(s->ring_end[k], st));
    for (ReqHost *r : L.finishers) SRV_HIP(hipEventRecord(r->ev_last, st));
    return std::string();
}

// Fill ring entry L.ring from the plan (caller holds the lock; the entry's
// previous step has finished, so its pinned tables are free).
void fill_step_tables(cms_server *s, StepLaunch &L) {
    const int k = L.ring;
    const SchedStep &p = L.plan;
    for (int slot = 0; slot < s->opts.slots; slot++) {
        float *o = nullptr;
        if (p.slot_owner[slot] >= 0) {
            const SchedRequest *sr = s->sched->get(p.slot_owner[slot]);
            o = (float *)(s->arena + sr->out_off) + (int64_t)p.slot_query[slot] * s->n_ref;
        }
        s->h_slot_out[k][slot] = o;
    }
    L.nqa = (int)p.active_blocks.size();
    for (int i = 0; i < L.nqa; i++) s->h_blocks[k][i] = p.active_blocks[i];
    L.ndirty = (int)p.dirty_blocks.size();
    for (int i = 0; i < L.ndirty; i++) {
        const int b = p.dirty_blocks[i];
        uint32_t m = 0;
        for (int j = 0; j < 32; j++)
            if (p.slot_owner[b * 32 + j] >= 0) m |= 1u << j;
        s->h_dirty[k][i] = b;
        s->h_masks[k][i] = m;
    }
}

void fatal_locked(cms_server *s, const std::string &msg) {
    if (s->fatal.empty()) s->fatal = msg;
    // every request not yet delivered fails with the device error
    for (auto &kv : s->reqs)
        if (!kv.second.finished) finish_locked(s, kv.first, CMS_E_HIP, "serving stopped: " + s->fatal);
    s->stop = true;
    s->cv_deliver.notify_all();
    s->cv_work.notify_all();
}

void worker_main(cms_server *s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (s->fatal.empty()) {
        bool progress = false;
        // 1. admission
        const std::vector<int64_t> adm = s->sched->admit();
        if (!adm.empty()) {
            progress = true;
            s->arena_peak = std::max(s->arena_peak, s->sched->arena().used());
            s->slots_peak = std::max(s->slots_peak, s->opts.slots - s->sched->free_slots());
            std::vector<std::pair<ReqHost *, int64_t>> jobs;
            for (int64_t id : adm) {
                ReqHost &r = s->reqs[id];
                r.t_admit = s->now();
                const SchedRequest *sr = s->sched->get(id);
                r.slots = sr->slots;
                jobs.push_back({&r, sr->scratch_off});
            }
            lk.unlock();
            std::vector<std::string> errs;
            for (auto &j : jobs) errs.push_back(launch_embedding(s, *j.first, j.second));
            lk.lock();
            for (size_t i = 0; i < adm.size(); i++) {
                if (errs[i].empty()) {
                    s->embedding_ids.push_back(adm[i]);
                } else {
                    fatal_locked(s, errs[i]);
                    break;
                }
            }
            if (!s->fatal.empty()) break;
        }
        // 2. embeddings finished -> ready to join
        for (size_t i = 0; i < s->embedding_ids.size();) {
            const int64_t id = s->embedding_ids[i];
            auto it = s->reqs.find(id);
            if (it == s->reqs.end() || it->second.finished) {
                s->embedding_ids.erase(s->embedding_ids.begin() + i);
                continue;
            }
            const hipError_t q = hipEventQuery(it->second.ev_embedded);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            if (q != hipSuccess) {
                fatal_locked(s, hip_msg("embedding", q));
                break;
            }
            s->sched->embedded(id);
            s->embedding_ids.erase(s->embedding_ids.begin() + i);
            progress = true;
        }
        if (!s->fatal.empty()) break;
        // 3. packed sets release their embedding scratch (a set delivered
        //    meanwhile released it with its output block)
        for (size_t i = 0; i < s->packing_ids.size();) {
            const int64_t id = s->packing_ids[i];
            auto it = s->reqs.find(id);
            if (it == s->reqs.end() || it->second.finished) {
                s->packing_ids.erase(s->packing_ids.begin() + i);
                continue;
            }
            const hipError_t q = hipEventQuery(it->second.ev_packed);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            if (q != hipSuccess) {
                fatal_locked(s, hip_msg("packing", q));
                break;
            }
            s->sched->scratch_released(id);
            s->packing_ids.erase(s->packing_ids.begin() + i);
            progress = true;
        }
        if (!s->fatal.empty()) break;
        // 4. retire finished steps
        while (!s->inflight.empty()) {
            const int k = s->inflight.front();
            const hipError_t q = hipEventQuery(s->ring_end[k]);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) {
                fatal_locked(s, hip_msg("step", q));
                break;
            }
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, s->ring_begin[k], s->ring_end[k]) == hipSuccess) s->kernel_ms += ms;
            s->inflight.pop_front();
            const double t = s->now();
            for (int64_t id : s->inflight_finishers.front()) {
                auto it = s->reqs.find(id);
                if (it != s->reqs.end() && !it->second.finished) it->second.t_computed = t;
            }
            s->inflight_finishers.pop_front();
            progress = true;
        }
        if (!s->fatal.empty()) break;
        // 5. plan and queue the next step
        if ((int)s->inflight.size() < kMaxInFlight) {
            StepLaunch L;
            if (s->sched->plan_step(L.plan)) {
                L.ring = (int)(L.plan.index % kRing);
                fill_step_tables(s, L);
                const double t = s->now();
                for (int64_t id : L.plan.joiners) {
                    ReqHost &r = s->reqs[id];
                    r.t_join = t;
                    r.join_step = L.plan.index;
                    L.joins.push_back({id, &r, s->sched->get(id)->scratch_off});
                }
                for (int64_t id : L.plan.finishers) L.finishers.push_back(&s->reqs[id]);
                lk.unlock();
                const std::string err = launch_step(s, L);
                lk.lock();
                if (!err.empty()) {
                    fatal_locked(s, err);
                    break;
                }
                s->steps++;
                s->inflight.push_back(L.ring);
                s->inflight_finishers.push_back(L.plan.finishers);
                for (auto &j : L.joins) s->packing_ids.push_back(j.id);
                for (int64_t id : L.plan.finishers) s->deliver_q.push_back(id);
                if (!L.plan.finishers.empty()) s->cv_deliver.notify_one();
                progress = true;
            }
        }
        const bool busy = !s->inflight.empty() || !s->embedding_ids.empty() || !s->packing_ids.empty() ||
                          s->sched->queued() > 0 || s->sched->active() > 0;
        if (s->stop && !busy) break;
        if (!progress) {
            if (busy)
                s->cv_work.wait_for(lk, std::chrono::microseconds(100));
            else
                s->cv_work.wait(lk);
        }
    }
    s->stop_delivery = true;
    s->cv_deliver.notify_all();
}

// Device arena -> caller's host buffer. Pinned (or registered) memory takes
// one DMA; pageable memory goes through the two pinned staging buffers: the
// DMA of piece i + 1 overlaps the parallel host copy of piece i.
hipError_t deliver(cms_server *s, float *dst, const uint8_t *src, size_t bytes) {
    hipPointerAttribute_t attr{};
    const bool pinned = hipPointerGetAttributes(&attr, dst) == hipSuccess && attr.type == hipMemoryTypeHost;
    (void)hipGetLastError();  // pageable memory is reported as an error; do not leave it pending
    hipError_t e;
    if (pinned || bytes == 0) {
        e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->s_copy);
        return e == hipSuccess ? hipStreamSynchronize(s->s_copy) : e;
    }
    const size_t piece = cms_server::kStageBytes;
    const size_t n = (bytes + piece - 1) / piece;
    auto issue = [&](size_t i) {
        const size_t len = std::min(piece, bytes - i * piece);
        hipError_t r = hipMemcpyAsync(s->stage[i & 1], src + i * piece, len, hipMemcpyDeviceToHost, s->s_copy);
        return r == hipSuccess ? hipEventRecord(s->ev_stage[i & 1], s->s_copy) : r;
    };
    e = issue(0);
    const int threads = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    for (size_t i = 0; i < n && e == hipSuccess; i++) {
        // staging buffer (i + 1) & 1 was drained by the host copy of piece i - 1
        if (i + 1 < n) e = issue(i + 1);
        if (e == hipSuccess) e = hipEventSynchronize(s->ev_stage[i & 1]);
        if (e == hipSuccess)
            parallel_copy(reinterpret_cast<uint8_t *>(dst) + i * piece, s->stage[i & 1],
                          std::min(piece, bytes - i * piece), threads);
    }
    if (e != hipSuccess) (void)hipStreamSynchronize(s->s_copy);
    return e;
}

void delivery_main(cms_server *s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_deliver.wait(lk, [&] { return !s->deliver_q.empty() || s->stop_delivery; });
        if (s->deliver_q.empty()) break;
        const int64_t id = s->deliver_q.front();
        s->deliver_q.pop_front();
        ReqHost &r = s->reqs[id];
        if (r.finished) continue;  // failed by a device error meanwhile
        const SchedRequest *sr = s->sched->get(id);
        const uint8_t *src = s->arena + sr->out_off;
        const size_t bytes = sizeof(float) * (size_t)r.nq * (size_t)s->n_ref;
        hipEvent_t ev = r.ev_last;
        float *dst = r.out;
        lk.unlock();
        hipError_t e = hipEventSynchronize(ev);
        if (e == hipSuccess) e = deliver(s, dst, src, bytes);
        lk.lock();
        if (e != hipSuccess) {
            fatal_locked(s, hip_msg("delivery", e));
            continue;
        }
        s->sched->delivered(id);
        finish_locked(s, id, CMS_OK, std::string());
        s->cv_work.notify_one();
    }
}

void free_server(cms_server *s) {
    (void)hipSetDevice(s->device);
    for (auto &kv : s->reqs) {
        ReqHost &r = kv.second;
        if (r.ev_embedded) (void)hipEventDestroy(r.ev_embedded);
        if (r.ev_packed) (void)hipEventDestroy(r.ev_packed);
        if (r.ev_last) (void)hipEventDestroy(r.ev_last);
    }
    s->reqs.clear();
    for (int k = 0; k < kRing; k++) {
        if (s->d_slot_out[k]) (void)hipFree(s->d_slot_out[k]);
        if (s->d_blocks[k]) (void)hipFree(s->d_blocks[k]);
        if (s->d_dirty[k]) (void)hipFree(s->d_dirty[k]);
        if (s->d_masks[k]) (void)hipFree(s->d_masks[k]);
        if (s->h_slot_out[k]) (void)hipHostFree(s->h_slot_out[k]);
        if (s->h_blocks[k]) (void)hipHostFree(s->h_blocks[k]);
        if (s->h_dirty[k]) (void)hipHostFree(s->h_dirty[k]);
        if (s->h_masks[k]) (void)hipHostFree(s->h_masks[k]);
        if (s->ring_begin[k]) (void)hipEventDestroy(s->ring_begin[k]);
        if (s->ring_end[k]) (void)hipEventDestroy(s->ring_end[k]);
    }
    if (s->arena) (void)hipFree(s->arena);
    if (s->QY) (void)hipFree(s->QY);
    if (s->qbm) (void)hipFree(s->qbm);
    if (s->qlbm) (void)hipFree(s->qlbm);
    if (s->qpl) (void)hipFree(s->qpl);
    if (s->qcrl) (void)hipFree(s->qcrl);
    if (s->qcolsum) (void)hipFree(s->qcolsum);
    if (s->work) (void)hipFree(s->work);
    if (s->counter) (void)hipFree(s->counter);
    if (s->part_acc) (void)hipFree(s->part_acc);
    if (s->s_embed) (void)hipStreamDestroy(s->s_embed);
    if (s->s_comp) (void)hipStreamDestroy(s->s_comp);
    if (s->s_copy) (void)hipStreamDestroy(s->s_copy);
    for (int k = 0; k < 2; k++) {
        if (s->stage[k]) (void)hipHostFree(s->stage[k]);
        if (s->ev_stage[k]) (void)hipEventDestroy(s->ev_stage[k]);
    }
    if (s->ctx) s->ctx->servers--;
    delete s;
}

#define CREATE_HIP(expr)                                                                               \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            free_server(s);                                                                            \
            return sfail(e_ == hipErrorOutOfMemory ? CMS_E_NOMEM : CMS_E_HIP, hip_msg(#expr, e_));     \
        }                                                                                              \
    } while (0)

}  // namespace

extern "C" {

void cms_server_opts_default(cms_server_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->slots = 1024;
    o->bands = 16;
}

int cms_server_create(cms_ctx *ref, const cms_server_opts *o, cms_server **out) {
    if (!ref || !out) return sfail(CMS_E_INVAL, "cms_server_create: null argument");
    *out = nullptr;
    cms_server_opts opts;
    cms_server_opts_default(&opts);
    if (o) opts = *o;
    if (opts.slots <= 0) opts.slots = 1024;
    if (opts.bands <= 0) opts.bands = 16;
    if (opts.slots % kSmBlock) return sfail(CMS_E_INVAL, "cms_server_create: slots must be a multiple of 32");
    if (opts.memory_budget < 0) return sfail(CMS_E_INVAL, "cms_server_create: negative memory budget");
    if (opts.variant < 0 || opts.variant > kSmCrossVariants)
        return sfail(CMS_E_INVAL, "cms_server_create: bad variant (the cross kernel has variants 1..8)");
    int rc = ctx_prepare(ref);
    if (rc) return rc;
    if (ref->n <= 0) return ctx_fail(ref, CMS_E_STATE, "cms_server_create: no reference samples");
    ctx_release_walk_layout(ref);  // queries are laid out in the server's own arena

    cms_server *s = new cms_server();
    s->ctx = ref;
    ref->servers++;
    s->device = ref->device;
    s->opts = opts;
    s->t0 = std::chrono::steady_clock::now();
    s->metric = ref->sm_metric;
    s->embnodes = s->metric == CMS_METRIC_UNWEIGHTED ? ref->d_embnodes_u : ref->d_embnodes;
    s->toprecs = s->metric == CMS_METRIC_UNWEIGHTED ? ref->d_toprecs_u : ref->d_toprecs;
    const SmData d = ctx_sm_data(ref);
    s->nrows = d.nrows;
    s->nwords = d.nwords;
    s->nlp = (int64_t)ref->tree.leaf_stream.size();
    s->nseg = (int64_t)ref->tree.seg_first.size();
    s->n_ref = ref->n;
    s->nrb = ((int64_t)ref->n + kSmBlock - 1) / kSmBlock;
    s->band_blocks = (s->nrb + opts.bands - 1) / opts.bands;
    s->scale_bits = ref->scale_bits;
    s->scale = std::ldexp(1.0, ref->scale_bits);
    // a query entry y(e) <= its column sum (weighted: the sample's consumed
    // leaf mass); fixed point must stay below kSmFixedMax as for the references
    s->mass_limit = (kSmFixedMax - 0.5 * (double)s->nrows - 64.0) / s->scale;
    if (s->metric == CMS_METRIC_UNWEIGHTED) {
        double wmax = 0.0;
        for (const WalkNode &n : ref->tree.nodes) wmax = std::max({wmax, (double)n.w1, (double)n.w2});
        if (wmax * s->scale + 0.5 >= kSmFixedMax) {
            free_server(s);
            return sfail(CMS_E_RANGE, "unweighted edge coefficient outside the references' fixed-point range");
        }
    }
    // default: split tiles (8-row chunk, one wave per workgroup) when the
    // references carry light records, else the plain (16, 4)
    const bool ref_light = d.lbm && d.crl && d.croff && d.pl && d.zrow >= d.nrows;
    s->variant = opts.variant > 0 ? opts.variant - 1 : (ref_light ? 6 : 1);
    if (s->variant >= 6 && !ref_light) {
        free_server(s);
        return sfail(CMS_E_INVAL, "cms_server_create: the split cross tiles need the references' light records");
    }
    CREATE_HIP(hipSetDevice(s->device));
    int cus = 0;
    CREATE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
    s->grid_waves = std::max(cus, 1) * 32;

    int64_t budget = opts.memory_budget;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        CREATE_HIP(hipMemGetInfo(&fr, &tot));
        budget = std::min<int64_t>(kDefaultBudget, (int64_t)(fr / 2));
    }
    s->opts.memory_budget = budget;
    s->sched = std::make_unique<Scheduler>(opts.slots, budget, opts.bands);
    CREATE_HIP(hipMalloc((void **)&s->arena, (size_t)std::max<int64_t>(budget, 256)));
    // QY: the entries' rows plus an all-zero row (row nrows, never packed),
    // which the split tiles stage in place of missing rows as they do Y's
    const size_t nqy = (size_t)(std::max<int64_t>(s->nrows, 0) + 1) * opts.slots;
    CREATE_HIP(hipMalloc((void **)&s->QY, sizeof(uint32_t) * nqy));
    CREATE_HIP(hipMemset(s->QY, 0, sizeof(uint32_t) * nqy));
    const size_t nqbm = (size_t)(opts.slots / kSmBlock) * (size_t)std::max<int64_t>(s->nwords, 1);
    CREATE_HIP(hipMalloc((void **)&s->qbm, sizeof(uint32_t) * nqbm));
    CREATE_HIP(hipMemset(s->qbm, 0, sizeof(uint32_t) * nqbm));
    s->qcrs = std::max<int64_t>(s->nwords, 1) * 32;
    if (s->variant >= 6) {
        CREATE_HIP(hipMalloc((void **)&s->qlbm, sizeof(uint32_t) * nqbm));
        CREATE_HIP(hipMemset(s->qlbm, 0, sizeof(uint32_t) * nqbm));
        CREATE_HIP(hipMalloc((void **)&s->qpl, sizeof(uint32_t) * nqbm));
        CREATE_HIP(hipMemset(s->qpl, 0, sizeof(uint32_t) * nqbm));
        CREATE_HIP(hipMalloc((void **)&s->qcrl, sizeof(uint4) * (size_t)(opts.slots / kSmBlock) * (size_t)s->qcrs));
    }
    CREATE_HIP(hipMalloc((void **)&s->qcolsum, sizeof(unsigned long long) * opts.slots));
    CREATE_HIP(hipMemset(s->qcolsum, 0, sizeof(unsigned long long) * opts.slots));
    CREATE_HIP(hipMalloc((void **)&s->work, sizeof(unsigned long long) * kWorkWords));
    CREATE_HIP(hipMalloc((void **)&s->counter, sizeof(unsigned long long) * kCounters));
    CREATE_HIP(hipMemset(s->counter, 0, sizeof(unsigned long long) * kCounters));
    s->part_tiles = (int64_t)(opts.slots / kSmBlock) * std::max<int64_t>(s->band_blocks, 1);
    CREATE_HIP(hipMalloc((void **)&s->part_acc, sizeof(uint32_t) * 1024 * (size_t)s->part_tiles));
    CREATE_HIP(hipMemset(s->part_acc, 0, sizeof(uint32_t) * 1024 * (size_t)s->part_tiles));
    const int nblk = opts.slots / kSmBlock;
    for (int k = 0; k < kRing; k++) {
        CREATE_HIP(hipMalloc((void **)&s->d_slot_out[k], sizeof(float *) * opts.slots));
        CREATE_HIP(hipMalloc((void **)&s->d_blocks[k], sizeof(int32_t) * nblk));
        CREATE_HIP(hipMalloc((void **)&s->d_dirty[k], sizeof(int32_t) * nblk));
        CREATE_HIP(hipMalloc((void **)&s->d_masks[k], sizeof(uint32_t) * nblk));
        CREATE_HIP(hipHostMalloc((void **)&s->h_slot_out[k], sizeof(float *) * opts.slots, hipHostMallocDefault));
        CREATE_HIP(hipHostMalloc((void **)&s->h_blocks[k], sizeof(int32_t) * nblk, hipHostMallocDefault));
        CREATE_HIP(hipHostMalloc((void **)&s->h_dirty[k], sizeof(int32_t) * nblk, hipHostMallocDefault));
        CREATE_HIP(hipHostMalloc((void **)&s->h_masks[k], sizeof(uint32_t) * nblk, hipHostMallocDefault));
        CREATE_HIP(hipEventCreate(&s->ring_begin[k]));
        CREATE_HIP(hipEventCreate(&s->ring_end[k]));
    }
    CREATE_HIP(hipStreamCreateWithFlags(&s->s_embed, hipStreamNonBlocking));
    CREATE_HIP(hipStreamCreateWithFlags(&s->s_comp, hipStreamNonBlocking));
    CREATE_HIP(hipStreamCreateWithFlags(&s->s_copy, hipStreamNonBlocking));
    for (int k = 0; k < 2; k++) {
        CREATE_HIP(hipHostMalloc((void **)&s->stage[k], cms_server::kStageBytes, hipHostMallocDefault));
        CREATE_HIP(hipEventCreateWithFlags(&s->ev_stage[k], hipEventDisableTiming));
    }
    // the references' layout was built on the context stream
    CREATE_HIP(hipStreamSynchronize(ref->stream));

    CrossData &x = s->cross;
    x.QY = s->QY;
    x.qstride = opts.slots;
    x.qbm = s->qbm;
    x.qcolsum = s->qcolsum;
    x.Y = d.Y;
    x.bm = d.bm;
    x.colsum = d.colsum;
    x.perm = d.perm;
    x.npad = d.npad;
    x.nwords = d.nwords;
    x.n_ref = d.n;
    x.unweighted = d.unweighted;
    x.inv_scale = d.inv_scale;
    x.counter = s->counter;
    x.qlbm = s->qlbm;
    x.qpl = s->qpl;
    x.qcrl = s->qcrl;
    x.qcrs = s->qcrs;
    x.lbm = d.lbm;
    x.pl = d.pl;
    x.crl = d.crl;
    x.croff = d.croff;
    x.zrow = d.zrow;  // the references' zero row; QY's is the same row index

    s->worker = std::thread(worker_main, s);
    s->deliverer = std::thread(delivery_main, s);
    *out = s;
    return CMS_OK;
}

int cms_server_submit(cms_server *s, int nq, const int64_t *rowptr, const int32_t *leaf, const float *val,
                      float *out, int64_t *id) {
    if (!s || !rowptr || !out || !id) return sfail(CMS_E_INVAL, "cms_server_submit: null argument");
    if (nq <= 0) return sfail(CMS_E_INVAL, "cms_server_submit: nq must be positive");
    if (rowptr[0] != 0) return sfail(CMS_E_INVAL, "cms_server_submit: rowptr[0] must be 0");
    for (int q = 0; q < nq; q++)
        if (rowptr[q + 1] < rowptr[q]) return sfail(CMS_E_INVAL, "cms_server_submit: rowptr must be non-decreasing");
    const int64_t nnz = rowptr[nq];
    if (nnz > 0 && (!leaf || !val)) return sfail(CMS_E_INVAL, "cms_server_submit: null leaf/val arrays");
    const cms_ctx *c = s->ctx;
    const int L = c->tree.leaf_n;
    for (int q = 0; q < nq; q++) {
        double mass = 0.0;
        for (int64_t k = rowptr[q]; k < rowptr[q + 1]; k++) {
            if (leaf[k] < 0 || leaf[k] >= L) return sfail(CMS_E_INVAL, "cms_server_submit: leaf id out of range");
            if (!(val[k] >= 0.0f) || !std::isfinite(val[k]))
                return sfail(CMS_E_UNSUPPORTED, "cms_server_submit: abundances must be finite and non-negative");
            mass += (double)val[k] * c->h_leaf_uses[leaf[k]];
        }
        if (s->metric == CMS_METRIC_WEIGHTED && mass > s->mass_limit)
            return sfail(CMS_E_RANGE, "cms_server_submit: query " + std::to_string(q) + " has mass " +
                                          std::to_string(mass) + ", beyond the references' fixed-point range " +
                                          std::to_string(s->mass_limit));
    }
    const ScratchLayout SL = scratch_layout(nq, nnz, s->nlp, s->nseg, s->nrows, s->nwords);
    const int64_t out_bytes = (int64_t)sizeof(float) * nq * (int64_t)s->n_ref;
    ReqHost r;
    r.nq = nq;
    r.nnz = nnz;
    r.out = out;
    r.rowptr.assign(rowptr, rowptr + nq + 1);
    if (nnz > 0) {
        r.col.assign(leaf, leaf + nnz);
        r.val.assign(val, val + nnz);
    }
    if (hipSetDevice(s->device) != hipSuccess || hipEventCreateWithFlags(&r.ev_embedded, hipEventDisableTiming) ||
        hipEventCreateWithFlags(&r.ev_packed, hipEventDisableTiming) ||
        hipEventCreateWithFlags(&r.ev_last, hipEventDisableTiming)) {
        if (r.ev_embedded) (void)hipEventDestroy(r.ev_embedded);
        if (r.ev_packed) (void)hipEventDestroy(r.ev_packed);
        return sfail(CMS_E_HIP, "cms_server_submit: event creation failed");
    }
    std::lock_guard<std::mutex> g(s->mu);
    auto destroy_events = [&]() {
        (void)hipEventDestroy(r.ev_embedded);
        (void)hipEventDestroy(r.ev_packed);
        (void)hipEventDestroy(r.ev_last);
    };
    if (!s->fatal.empty()) {
        destroy_events();
        return sfail(CMS_E_HIP, "serving stopped: " + s->fatal);
    }
    if (s->stop) {
        destroy_events();
        return sfail(CMS_E_STATE, "cms_server_submit: the server is shutting down");
    }
    // densest queries first into the set's slots: 32-slot blocks then hold
    // queries of similar sparsity, whose root-path union (the rows a tile
    // visits) stays small for the sparse blocks
    std::vector<int32_t> order((size_t)nq);
    for (int q = 0; q < nq; q++) order[(size_t)q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return rowptr[a + 1] - rowptr[a] > rowptr[b + 1] - rowptr[b];
    });
    const int64_t rid = s->next_id++;
    std::string err;
    if (!s->sched->submit(rid, nq, SL.total, out_bytes, err, order)) {
        destroy_events();
        return sfail(nq > s->opts.slots ? CMS_E_INVAL : CMS_E_NOMEM, "cms_server_submit: " + err);
    }
    r.t_submit = s->now();
    s->reqs.emplace(rid, std::move(r));
    s->submitted++;
    *id = rid;
    s->cv_work.notify_one();
    return CMS_OK;
}

int cms_server_wait(cms_server *s, int64_t id, double timeout_s, cms_request_info *info) {
    if (!s) return sfail(CMS_E_INVAL, "cms_server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return sfail(CMS_E_INVAL, "cms_server_wait: unknown request id");
    auto done = [&] { return s->reqs[id].finished; };
    if (timeout_s < 0)
        s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
        return sfail(CMS_E_TIMEOUT, "cms_server_wait: request still running");
    ReqHost &r = s->reqs[id];
    if (info) {
        info->rc = r.rc;
        info->nq = r.nq;
        info->join_step = r.join_step;
        info->t_submit = r.t_submit;
        info->t_admit = r.t_admit;
        info->t_join = r.t_join;
        info->t_done = r.t_done;
        info->t_computed = r.t_computed;
    }
    const int rc = r.rc;
    if (rc != CMS_OK) set_global_error(r.err);
    (void)hipEventDestroy(r.ev_embedded);
    (void)hipEventDestroy(r.ev_packed);
    (void)hipEventDestroy(r.ev_last);
    s->reqs.erase(id);
    s->sched->forget(id);
    return rc;
}

int cms_server_drain(cms_server *s, double timeout_s) {
    if (!s) return sfail(CMS_E_INVAL, "cms_server_drain: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto all_done = [&] {
        for (auto &kv : s->reqs)
            if (!kv.second.finished) return false;
        return true;
    };
    if (timeout_s < 0)
        s->cv_done.wait(lk, all_done);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), all_done))
        return sfail(CMS_E_TIMEOUT, "cms_server_drain: requests still running");
    if (!s->fatal.empty()) return sfail(CMS_E_HIP, "serving stopped: " + s->fatal);
    return CMS_OK;
}

int cms_server_get_stats(cms_server *s, cms_server_stats *st) {
    if (!s || !st) return sfail(CMS_E_INVAL, "cms_server_get_stats: null argument");
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->nfailed;
    st->queued = (int64_t)s->sched->queued();
    st->active = (int64_t)s->sched->active();
    st->steps = s->steps;
    st->kernel_ms = s->kernel_ms;
    st->pairs = s->pairs;
    unsigned long long cnt[kCounters] = {};
    if (s->counter && hipMemcpy(cnt, s->counter, sizeof(cnt), hipMemcpyDeviceToHost) == hipSuccess) {
        st->pair_entries = (double)cnt[kCntVisited] * 1024.0;
        st->entries_staged = (double)cnt[kCntStaged];
        st->entries_sparse = (double)cnt[kCntSparseRows];
        st->sparse_pairs = (double)cnt[kCntSparsePairs];
        st->entries_half = (double)cnt[kCntHalfRows];
        st->half_pairs = (double)cnt[kCntHalfPairs];
    }
    st->arena_bytes = s->opts.memory_budget;
    st->arena_peak = s->arena_peak;
    st->slots = s->opts.slots;
    st->slots_peak = s->slots_peak;
    st->bands = s->opts.bands;
    st->uptime_s = s->now();
    return CMS_OK;
}

void cms_server_destroy(cms_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    if (s->deliverer.joinable()) s->deliverer.join();
    (void)hipSetDevice(s->device);
    if (s->s_comp) (void)hipStreamSynchronize(s->s_comp);
    if (s->s_embed) (void)hipStreamSynchronize(s->s_embed);
    if (s->s_copy) (void)hipStreamSynchronize(s->s_copy);
    free_server(s);
}

}  // extern "C"
