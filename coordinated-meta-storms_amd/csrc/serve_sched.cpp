// Serving scheduler: arena, slots, FIFO admission and the band cycle.
#include "serve_sched.h"

#include <algorithm>
#include <cstring>

namespace cms {

// ------------------------------------------------------------------ Arena

Arena::Arena(int64_t capacity, int64_t align) : capacity_(capacity > 0 ? capacity : 0), align_(align > 0 ? align : 1) {
    if (capacity_ > 0) free_[0] = capacity_;
}

int64_t Arena::allocate(int64_t bytes) {
    const int64_t need = std::max<int64_t>(1, (bytes + align_ - 1) / align_) * align_;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < need) continue;
        const int64_t off = it->first, rest = it->second - need;
        free_.erase(it);
        if (rest > 0) free_[off + need] = rest;
        live_[off] = need;
        used_ += need;
        return off;
    }
    return -1;
}

void Arena::release(int64_t offset) {
    auto lv = live_.find(offset);
    if (lv == live_.end()) return;
    int64_t off = offset, size = lv->second;
    used_ -= size;
    live_.erase(lv);
    auto next = free_.lower_bound(off);
    if (next != free_.end() && next->first == off + size) {
        size += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += size;
            return;
        }
    }
    free_[off] = size;
}

int64_t Arena::largest_free() const {
    int64_t m = 0;
    for (const auto &f : free_) m = std::max(m, f.second);
    return m;
}

bool Arena::check() const {
    // free and live blocks tile [0, capacity) without overlap; free blocks
    // never touch each other (they would have been merged)
    std::map<int64_t, std::pair<int64_t, bool>> all;
    for (const auto &f : free_) all[f.first] = {f.second, true};
    for (const auto &l : live_) {
        if (all.count(l.first)) return false;
        all[l.first] = {l.second, false};
    }
    int64_t at = 0, used = 0;
    bool prev_free = false;
    for (const auto &b : all) {
        if (b.first != at || b.second.first <= 0) return false;
        if (b.second.second && prev_free) return false;
        prev_free = b.second.second;
        if (!b.second.second) used += b.second.first;
        at += b.second.first;
    }
    return at == capacity_ && used == used_;
}

// -------------------------------------------------------------- Scheduler

Scheduler::Scheduler(int slots, int64_t arena_bytes, int bands)
    : nslots_(slots), bands_(std::max(bands, 1)), arena_(arena_bytes), owner_((size_t)std::max(slots, 0), -1),
      free_count_(std::max(slots, 0)), dirty_((size_t)(std::max(slots, 0) + 31) / 32, 0) {}

bool Scheduler::submit(int64_t id, int nq, int64_t scratch_bytes, int64_t out_bytes, std::string &err,
                       const std::vector<int32_t> &order) {
    if (nq <= 0 || nq > nslots_) {
        err = "query set of " + std::to_string(nq) + " samples: the server holds 1.." + std::to_string(nslots_);
        return false;
    }
    if (!order.empty()) {
        std::vector<char> seen((size_t)nq, 0);
        bool ok = (int)order.size() == nq;
        for (size_t i = 0; ok && i < order.size(); i++) {
            ok = order[i] >= 0 && order[i] < nq && !seen[(size_t)order[i]];
            if (ok) seen[(size_t)order[i]] = 1;
        }
        if (!ok) {
            err = "query placement order is not a permutation of the set";
            return false;
        }
    }
    Arena probe(arena_.capacity());
    if (probe.allocate(scratch_bytes) < 0 || probe.allocate(out_bytes) < 0) {
        err = "query set needs " + std::to_string(scratch_bytes + out_bytes) + " bytes of device memory; the budget is " +
              std::to_string(arena_.capacity());
        return false;
    }
    SchedRequest r;
    r.id = id;
    r.nq = nq;
    r.scratch_bytes = scratch_bytes;
    r.out_bytes = out_bytes;
    r.order = order;
    reqs_[id] = r;
    queue_.push_back(id);
    return true;
}

std::vector<int64_t> Scheduler::admit() {
    std::vector<int64_t> out;
    while (!queue_.empty()) {
        SchedRequest &r = reqs_[queue_.front()];
        if (r.nq > free_count_) break;
        const int64_t s = arena_.allocate(r.scratch_bytes);
        if (s < 0) break;
        const int64_t o = arena_.allocate(r.out_bytes);
        if (o < 0) {
            arena_.release(s);
            break;
        }
        r.scratch_off = s;
        r.out_off = o;
        r.slots.assign((size_t)r.nq, -1);
        int placed = 0;
        for (int k = 0; k < nslots_ && placed < r.nq; k++)
            if (owner_[k] < 0) {
                owner_[k] = r.id;
                const int q = r.order.empty() ? placed : r.order[(size_t)placed];
                r.slots[(size_t)q] = k;
                placed++;
            }
        free_count_ -= r.nq;
        r.state = ReqState::Embedding;
        out.push_back(r.id);
        queue_.pop_front();
    }
    return out;
}

void Scheduler::embedded(int64_t id) {
    auto it = reqs_.find(id);
    if (it == reqs_.end() || it->second.state != ReqState::Embedding) return;
    it->second.state = ReqState::Ready;
    ready_.push_back(id);
}

bool Scheduler::plan_step(SchedStep &step) {
    if (ready_.empty() && active_.empty()) return false;
    step = SchedStep();
    step.index = next_step_;
    step.band = (int)(next_step_ % bands_);
    for (int64_t id : ready_) {
        SchedRequest &r = reqs_[id];
        r.state = ReqState::Active;
        r.join_step = next_step_;
        r.steps_done = 0;
        active_.push_back(id);
        step.joiners.push_back(id);
        for (int32_t s : r.slots) dirty_[s / 32] = 1;
    }
    ready_.clear();
    step.slot_owner.assign((size_t)nslots_, -1);
    step.slot_query.assign((size_t)nslots_, -1);
    for (int64_t id : active_) {
        const SchedRequest &r = reqs_[id];
        for (int q = 0; q < r.nq; q++) {
            step.slot_owner[r.slots[q]] = id;
            step.slot_query[r.slots[q]] = q;
        }
    }
    const int nblk = (nslots_ + 31) / 32;
    for (int b = 0; b < nblk; b++) {
        bool any = false;
        for (int k = b * 32; k < std::min(nslots_, b * 32 + 32) && !any; k++) any = step.slot_owner[k] >= 0;
        if (any) step.active_blocks.push_back(b);
        if (dirty_[b]) {
            step.dirty_blocks.push_back(b);
            dirty_[b] = 0;
        }
    }
    std::vector<int64_t> still;
    for (int64_t id : active_) {
        SchedRequest &r = reqs_[id];
        if (++r.steps_done == bands_) {
            step.finishers.push_back(id);
            r.state = ReqState::Draining;
            free_slots_of(r);
        } else {
            still.push_back(id);
        }
    }
    active_.swap(still);
    next_step_++;
    return true;
}

void Scheduler::free_slots_of(SchedRequest &r) {
    for (int32_t s : r.slots) {
        if (owner_[s] == r.id) {
            owner_[s] = -1;
            free_count_++;
        }
        dirty_[s / 32] = 1;
    }
}

void Scheduler::scratch_released(int64_t id) {
    auto it = reqs_.find(id);
    if (it == reqs_.end() || it->second.scratch_off < 0) return;
    arena_.release(it->second.scratch_off);
    it->second.scratch_off = -1;
}

void Scheduler::delivered(int64_t id) {
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return;
    SchedRequest &r = it->second;
    if (r.scratch_off >= 0) arena_.release(r.scratch_off);
    if (r.out_off >= 0) arena_.release(r.out_off);
    r.scratch_off = r.out_off = -1;
    r.state = ReqState::Done;
}

void Scheduler::failed(int64_t id) {
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return;
    SchedRequest &r = it->second;
    if (r.state == ReqState::Queued) queue_.erase(std::find(queue_.begin(), queue_.end(), id));
    ready_.erase(std::remove(ready_.begin(), ready_.end(), id), ready_.end());
    if (r.state == ReqState::Embedding || r.state == ReqState::Ready) free_slots_of(r);
    if (r.scratch_off >= 0) arena_.release(r.scratch_off);
    if (r.out_off >= 0) arena_.release(r.out_off);
    r.scratch_off = r.out_off = -1;
    r.state = ReqState::Failed;
}

void Scheduler::forget(int64_t id) {
    auto it = reqs_.find(id);
    if (it != reqs_.end() && (it->second.state == ReqState::Done || it->second.state == ReqState::Failed))
        reqs_.erase(it);
}

const SchedRequest *Scheduler::get(int64_t id) const {
    auto it = reqs_.find(id);
    return it == reqs_.end() ? nullptr : &it->second;
}

size_t Scheduler::active() const { return active_.size(); }

void parallel_copy(void *dst, const void *src, size_t bytes, int threads) {
    constexpr size_t kPiece = (size_t)4 << 20;
    const int64_t pieces = (int64_t)((bytes + kPiece - 1) / kPiece);
    char *d = static_cast<char *>(dst);
    const char *s = static_cast<const char *>(src);
#pragma omp parallel for schedule(static) num_threads(threads) if (pieces > 1)
    for (int64_t k = 0; k < pieces; k++) {
        const size_t off = (size_t)k * kPiece;
        std::memcpy(d + off, s + off, std::min(kPiece, bytes - off));
    }
}

}  // namespace cms
