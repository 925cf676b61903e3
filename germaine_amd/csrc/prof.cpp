#include "prof.hpp"

#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/gfdfd.h"

namespace gfdfd {

namespace {
// Allocated once and never destroyed: device buffers held by other static
// objects (the plan cache) are freed during static destruction at exit, in
// an order across translation units that C++ leaves open.
struct MemBook {
    std::mutex mu;
    std::unordered_map<void *, size_t> live;
    size_t now = 0, peak = 0;
};
MemBook &mem_book() {
    static MemBook *b = new MemBook();
    return *b;
}
}  // namespace

void mem_track_alloc(void *p, size_t bytes) {
    MemBook &b = mem_book();
    std::lock_guard<std::mutex> g(b.mu);
    b.live[p] = bytes;
    b.now += bytes;
    if (b.now > b.peak) b.peak = b.now;
}

void mem_track_free(void *p) {
    MemBook &b = mem_book();
    std::lock_guard<std::mutex> g(b.mu);
    auto it = b.live.find(p);
    if (it == b.live.end()) return;
    b.now -= it->second;
    b.live.erase(it);
}

const char *const kClassNames[K_NCLASS] = {
    "assemble",   "front_assemble", "scatter_orig", "small_lu",   "panel",       "panel_apply", "trsm_u",  "gemm_update",
    "diag_inv",   "fwd_gather",     "fwd_leaf",     "fwd_small1",   "fwd_front1", "fwd_front",   "fwd_blk", "bwd_small1",
    "bwd_front1", "bwd_front",      "bwd_contrib",  "bwd_blk",    "piv_unswap",  "fwd_vec", "bwd_vec",
    "permute_in", "permute_out",    "residual",     "norm_reduce", "axpy",    "subtree"};
static_assert(sizeof(kClassNames) / sizeof(kClassNames[0]) == K_NCLASS, "one name per kernel class");

namespace {
struct Rec {
    KClass c;
    hipEvent_t a, b;
    double flops, bytes;
};
struct Totals {
    double ms = 0, flops = 0, bytes = 0;
    long long launches = 0;
};
std::mutex mu;
bool enabled = false;
bool selected[K_NCLASS] = {};   // classes that record events while enabled
bool select_all = true;
std::vector<Rec> pending;
std::vector<hipEvent_t> pool;
Totals totals[K_NCLASS];

hipEvent_t get_event() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void drain() {
    for (Rec &r : pending) {
        float ms = 0;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            Totals &t = totals[r.c];
            t.ms += ms;
            t.flops += r.flops;
            t.bytes += r.bytes;
            t.launches++;
        }
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    pending.clear();
}
}  // namespace

static thread_local int paused = 0;   // ProfPause depth of this thread

bool prof_enabled() { return enabled && paused == 0; }

ProfPause::ProfPause() { paused++; }
ProfPause::~ProfPause() { paused--; }

ProfScope::ProfScope(KClass c, hipStream_t st, double flops, double bytes) : st_(st) {
    if (!enabled || paused || !(select_all || selected[c])) return;
    std::lock_guard<std::mutex> lk(mu);
    Rec r{c, get_event(), get_event(), flops, bytes};
    if (!r.a || !r.b) return;
    if (hipEventRecord(r.a, st) != hipSuccess) return;
    pending.push_back(r);
    slot_ = (int)pending.size() - 1;
}

ProfScope::~ProfScope() {
    if (slot_ < 0) return;
    std::lock_guard<std::mutex> lk(mu);
    (void)hipEventRecord(pending[slot_].b, st_);
    if (pending.size() > 4096) drain();
}

}  // namespace gfdfd

using namespace gfdfd;

extern "C" {

int gfdfd_device_memory(long long *current_bytes, long long *peak_bytes, int reset_peak) {
    MemBook &b = mem_book();
    std::lock_guard<std::mutex> g(b.mu);
    if (current_bytes) *current_bytes = (long long)b.now;
    if (peak_bytes) *peak_bytes = (long long)b.peak;
    if (reset_peak) b.peak = b.now;
    return GFDFD_OK;
}

int gfdfd_profile(int enable) {
    std::lock_guard<std::mutex> lk(mu);
    if (!enable) drain();
    enabled = enable != 0;
    return GFDFD_OK;
}

int gfdfd_profile_select(const char *kernel) {
    std::lock_guard<std::mutex> lk(mu);
    if (!kernel || !*kernel) {
        select_all = true;
        return GFDFD_OK;
    }
    for (int c = 0; c < K_NCLASS; c++)
        if (std::strcmp(kernel, kClassNames[c]) == 0) {
            for (bool &b : selected) b = false;
            selected[c] = true;
            select_all = false;
            return GFDFD_OK;
        }
    return GFDFD_ERR_ARG;
}

void gfdfd_profile_reset(void) {
    std::lock_guard<std::mutex> lk(mu);
    drain();
    for (auto &t : totals) t = Totals();
}

int gfdfd_profile_read(const char *kernel, double *ms, long long *launches, double *flops, double *bytes) {
    std::lock_guard<std::mutex> lk(mu);
    drain();
    for (int c = 0; c < K_NCLASS; c++)
        if (kernel && std::strcmp(kernel, kClassNames[c]) == 0) {
            if (ms) *ms = totals[c].ms;
            if (launches) *launches = totals[c].launches;
            if (flops) *flops = totals[c].flops;
            if (bytes) *bytes = totals[c].bytes;
            return GFDFD_OK;
        }
    return GFDFD_ERR_ARG;
}

}  // extern "C"
