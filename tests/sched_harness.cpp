// Test driver of the serving scheduler (coordinated-meta-storms_amd/csrc/
// serve_sched.cpp), built by tests/test_serve_sched.py with g++: no GPU.
//
//   sched_harness arena SEED OPS      random allocate/release against a
//                                     brute-force byte map; prints "ok" or
//                                     the first mismatch
//   sched_harness sim < scenario      discrete-time run of the serving loop
//                                     policy, one step per tick; prints a
//                                     trace of events for the test to check
//
// Scenario: first line "slots arena bands embed_ticks deliver_ticks", then
// one line per query set "arrive_tick nq scratch_bytes out_bytes reverse"
// (reverse 1: the set's queries take its slots last query first).
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "serve_sched.h"

using cms::Arena;
using cms::SchedStep;
using cms::Scheduler;

static int arena_fuzz(unsigned seed, int ops) {
    const int64_t cap = 1 << 16;
    Arena a(cap, 64);
    std::vector<int> owner(cap, -1);  // byte -> block id
    std::map<int64_t, std::pair<int64_t, int>> live;  // offset -> (size, id)
    std::mt19937 rng(seed);
    int next = 0;
    for (int i = 0; i < ops; i++) {
        if (live.empty() || rng() % 3 != 0) {
            const int64_t want = 1 + rng() % 4000;
            const int64_t off = a.allocate(want);
            // brute force: is there an aligned free run of the rounded size?
            const int64_t need = (want + 63) / 64 * 64;
            bool exists = false;
            for (int64_t o = 0; o + need <= cap && !exists; o += 64) {
                bool free_run = true;
                for (int64_t b = o; b < o + need && free_run; b++) free_run = owner[b] < 0;
                exists = free_run;
            }
            if (off < 0) {
                // first fit over coalesced free ranges finds any aligned run
                if (exists) return printf("allocate %lld failed with room\n", (long long)want), 1;
                continue;
            }
            if (off % 64 || off + need > cap) return printf("bad offset %lld\n", (long long)off), 1;
            for (int64_t b = off; b < off + need; b++) {
                if (owner[b] >= 0) return printf("overlap at %lld\n", (long long)b), 1;
                owner[b] = next;
            }
            live[off] = {need, next++};
        } else {
            auto it = live.begin();
            std::advance(it, rng() % live.size());
            a.release(it->first);
            for (int64_t b = it->first; b < it->first + it->second.first; b++) owner[b] = -1;
            live.erase(it);
        }
        if (!a.check()) return printf("check failed at op %d\n", i), 1;
        int64_t used = 0;
        for (auto &l : live) used += l.second.first;
        if (used != a.used()) return printf("used mismatch\n"), 1;
    }
    for (auto &l : live) a.release(l.first);
    if (a.used() != 0 || !a.check() || a.largest_free() != cap) return printf("not empty at end\n"), 1;
    printf("ok\n");
    return 0;
}

static int sim() {
    int slots, bands, embed_ticks, deliver_ticks;
    long long arena;
    if (scanf("%d %lld %d %d %d", &slots, &arena, &bands, &embed_ticks, &deliver_ticks) != 5) return 2;
    struct Arrival {
        int tick, nq;
        long long scratch, out;
        int reverse;
    };
    std::vector<Arrival> arr;
    Arrival x;
    while (scanf("%d %d %lld %lld %d", &x.tick, &x.nq, &x.scratch, &x.out, &x.reverse) == 5) arr.push_back(x);
    Scheduler s(slots, arena, bands);
    std::map<int64_t, int> embed_due, deliver_due, scratch_due;
    size_t next_arrival = 0;
    int64_t done = 0, accepted = 0;
    for (int t = 0; t < 100000; t++) {
        while (next_arrival < arr.size() && arr[next_arrival].tick <= t) {
            const Arrival &a = arr[next_arrival];
            const int64_t id = (int64_t)next_arrival + 1;
            std::string err;
            std::vector<int32_t> order;
            if (a.reverse)
                for (int q = a.nq - 1; q >= 0; q--) order.push_back(q);
            if (s.submit(id, a.nq, a.scratch, a.out, err, order)) {
                printf("submit %d %lld ok\n", t, (long long)id);
                accepted++;
            } else {
                printf("submit %d %lld reject\n", t, (long long)id);
            }
            next_arrival++;
        }
        for (int64_t id : s.admit()) {
            printf("admit %d %lld", t, (long long)id);
            for (int32_t sl : s.get(id)->slots) printf(" %d", sl);
            printf("\n");
            embed_due[id] = t + embed_ticks;
        }
        for (auto it = embed_due.begin(); it != embed_due.end();) {
            if (it->second <= t) {
                s.embedded(it->first);
                it = embed_due.erase(it);
            } else {
                ++it;
            }
        }
        SchedStep st;
        if (s.plan_step(st)) {
            for (int64_t id : st.joiners) {
                printf("join %d %lld %lld\n", t, (long long)id, (long long)st.index);
                scratch_due[id] = t + 1;
            }
            printf("step %d %lld %d", t, (long long)st.index, st.band);
            std::map<int64_t, int> per;
            for (int k = 0; k < slots; k++)
                if (st.slot_owner[k] >= 0) per[st.slot_owner[k]]++;
            for (auto &p : per) printf(" %lld:%d", (long long)p.first, p.second);
            printf("\n");
            printf("blocks %d", t);
            for (int32_t b : st.active_blocks) printf(" %d", b);
            printf(" |");
            for (int32_t b : st.dirty_blocks) printf(" %d", b);
            printf("\n");
            for (int64_t id : st.finishers) {
                printf("finish %d %lld %lld\n", t, (long long)id, (long long)st.index);
                deliver_due[id] = t + deliver_ticks;
            }
        }
        for (auto it = scratch_due.begin(); it != scratch_due.end();) {
            if (it->second <= t) {
                s.scratch_released(it->first);
                it = scratch_due.erase(it);
            } else {
                ++it;
            }
        }
        for (auto it = deliver_due.begin(); it != deliver_due.end();) {
            if (it->second <= t) {
                s.delivered(it->first);
                printf("deliver %d %lld\n", t, (long long)it->first);
                s.forget(it->first);
                scratch_due.erase(it->first);
                it = deliver_due.erase(it);
                done++;
            } else {
                ++it;
            }
        }
        printf("arena %d %lld %d %d\n", t, (long long)s.arena().used(), s.arena().check() ? 1 : 0, s.free_slots());
        if (next_arrival == arr.size() && done == accepted) break;
    }
    printf("end %lld %lld\n", (long long)done, (long long)accepted);
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 4 && std::string(argv[1]) == "arena") return arena_fuzz((unsigned)atoi(argv[2]), atoi(argv[3]));
    if (argc >= 2 && std::string(argv[1]) == "sim") return sim();
    fprintf(stderr, "usage: sched_harness arena SEED OPS | sim < scenario\n");
    return 2;
}
