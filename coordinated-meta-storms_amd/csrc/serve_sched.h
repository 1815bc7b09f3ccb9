// Host-side scheduling state of the serving loop: device-arena allocation,
// query-slot allocation, FIFO admission and the band cycle that every
// admitted query set runs through. Pure C++ (no HIP), so the policy runs and
// is tested without a GPU; cms_serve.hip drives it and does the device work.
//
// Model. The resident reference samples are split into `bands` contiguous
// ranges of 32-sample blocks. The device runs one step at a time; a step
// compares every active query slot with one band. A query set that joins at
// step s is compared with bands s, s+1, ..., s+bands-1 (mod bands), i.e. each
// band exactly once, and leaves after its last step; query sets join and
// leave at any step boundary (continuous batching), so slots and memory are
// reused while other sets are mid-cycle.
#pragma once

#include <cstdint>
#include <deque>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

namespace cms {

// First-fit allocator over a byte range [0, capacity) with aligned blocks;
// freed blocks coalesce with their neighbours.
class Arena {
public:
    explicit Arena(int64_t capacity = 0, int64_t align = 256);
    // Offset of a new block of `bytes`, or -1 when no free range fits.
    int64_t allocate(int64_t bytes);
    void release(int64_t offset);
    int64_t capacity() const { return capacity_; }
    int64_t used() const { return used_; }
    int64_t largest_free() const;
    size_t blocks() const { return live_.size(); }
    // Consistency of the free list and the live blocks (tests).
    bool check() const;

private:
    int64_t capacity_, align_, used_ = 0;
    std::map<int64_t, int64_t> free_;  // offset -> size
    std::map<int64_t, int64_t> live_;  // offset -> size
};

enum class ReqState { Queued, Embedding, Ready, Active, Draining, Done, Failed };

struct SchedRequest {
    int64_t id = 0;
    int nq = 0;
    int64_t scratch_bytes = 0, out_bytes = 0;
    int64_t scratch_off = -1, out_off = -1;
    std::vector<int32_t> slots;  // query q -> slot
    std::vector<int32_t> order;  // placement order of the queries (empty = 0..nq-1)
    ReqState state = ReqState::Queued;
    int64_t join_step = -1;      // step at which it joined
    int steps_done = 0;
};

// One step of the cycle as planned by the scheduler.
struct SchedStep {
    int64_t index = 0;
    int band = 0;
    std::vector<int64_t> joiners;       // ids joining at this step (pack before the step)
    std::vector<int64_t> finishers;     // ids whose last band is this step
    std::vector<int32_t> active_blocks; // 32-slot blocks holding an active slot, ascending
    std::vector<int32_t> dirty_blocks;  // blocks whose slot membership changed since the last step
    std::vector<int64_t> slot_owner;    // per slot: request id active in this step, -1 = none
    std::vector<int32_t> slot_query;    // per slot: query index within its request
};

class Scheduler {
public:
    Scheduler(int slots, int64_t arena_bytes, int bands);

    // Queue a query set (FIFO). Returns false with `err` when it can never be
    // admitted: more queries than slots or more memory than the arena.
    // `order` (a permutation of 0..nq-1, or empty) is the order in which the
    // queries take the set's slots, lowest slot first: queries adjacent in it
    // share 32-slot blocks.
    bool submit(int64_t id, int nq, int64_t scratch_bytes, int64_t out_bytes, std::string &err,
                const std::vector<int32_t> &order = {});
    // Admit queued sets in FIFO order while the head fits (free slots and
    // both arena blocks); returns the ids admitted now (state Embedding).
    std::vector<int64_t> admit();
    // The set's queries are embedded: it joins at the next planned step.
    void embedded(int64_t id);
    // Plan the next step; false when no set is ready or active.
    bool plan_step(SchedStep &step);
    // Embedding scratch no longer read by the device (packed into slots).
    void scratch_released(int64_t id);
    // Output delivered: the output block is freed and the set is Done.
    void delivered(int64_t id);
    // Failure before joining (embedding error): frees its slots and memory.
    void failed(int64_t id);
    // Forget a Done / Failed set.
    void forget(int64_t id);

    const SchedRequest *get(int64_t id) const;
    int slots() const { return nslots_; }
    int bands() const { return bands_; }
    int free_slots() const { return free_count_; }
    int64_t steps_planned() const { return next_step_; }
    size_t queued() const { return queue_.size(); }
    size_t active() const;
    const Arena &arena() const { return arena_; }

private:
    void free_slots_of(SchedRequest &r);

    int nslots_, bands_;
    Arena arena_;
    std::vector<int64_t> owner_;   // slot -> id holding it (admitted or active), -1 free
    int free_count_;
    std::deque<int64_t> queue_;
    std::unordered_map<int64_t, SchedRequest> reqs_;
    std::vector<int64_t> ready_;   // ids embedded, joining at the next step (admission order)
    std::vector<int64_t> active_;  // ids in the cycle (join order)
    std::vector<char> dirty_;      // per block: membership changed since the last step
    int64_t next_step_ = 0;
};

// Host memcpy split over up to `threads` OpenMP threads (4 MiB pieces): the
// serving loop's delivery from its pinned staging buffers into callers'
// pageable output, where one thread's copy and first-touch page faults set
// the pace.
void parallel_copy(void *dst, const void *src, size_t bytes, int threads);

}  // namespace cms
