// Internal declarations of libsi.so (not installed; the public ABI is
// include/set_intersection.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "set_intersection.h"

namespace si_impl {

// ---------------------------------------------------------------- errors --
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);

#define SI_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return ::si_impl::fail(SI_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SI_TRY(expr)          \
    do {                      \
        int s_ = (expr);      \
        if (s_ != SI_OK) return s_; \
    } while (0)

// Device memory owned by the library, grown on demand.
struct DevMem {
    void* ptr = nullptr;
    uint64_t bytes = 0;
    int ensure(uint64_t want);  // grows (never shrinks); contents are not kept
    void release();
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

struct BatchPlan;
void free_batch_plan(BatchPlan* p);

}  // namespace si_impl

struct si_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for kernels that run beside the
    // main stream's; joined back to it through the two events
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // copy stream (default priority) of si_buffer_download_async, ordered
    // after the context stream through copy_fork
    hipStream_t copy = nullptr;
    hipEvent_t copy_fork = nullptr;
    std::set<void*> host_allocs;  // page-locked host memory from si_host_alloc
    hipDeviceProp_t prop{};
    int num_cus = 0;
    // DeviceTimer events
    std::vector<hipEvent_t> events;
    // per-kernel timing
    bool kernel_timing = false;
    struct Pending {
        std::string name;
        hipEvent_t start, end;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    struct KStat {
        double ms = 0.0;
        uint64_t launches = 0;
    };
    std::map<std::string, KStat> kstats;
    // scratch workspaces (reused across calls on this context)
    si_impl::DevMem ws_small;   // counters, per-segment counts, reduction partials
    si_impl::DevMem ws_stage;   // materialisation staging
    si_impl::DevMem ws_lookback;  // SISI materialisation look-back states (zero between launches)
    si_impl::DevMem ws_aux;     // technique-specific (bitmaps, hash bins, pair lists)
    si_impl::DevMem ws_aux2;
    bool ws_small_zeroed = false;
    int sisi_shape = 0;  // tile shape of the SISI merge kernel used by materialisation
    int sisi_prio = 1;   // SISI merge kernel rotates wave priority by tile (SI_SISI_PRIO=0 turns it off)
    int sisi_ramp = 4;   // materialising segments grow by 2 ramp/256 of the mean from first to last (SI_SISI_RAMP)
    // si_pair_count_noalloc: this device's copy of the file-scope scratch
    // array, and whether the current call must take its workspace from it
    void* noalloc_ws = nullptr;
    bool in_noalloc = false;
    // objects created from this context; si_close frees whatever is left
    std::set<si_buffer*> buffers;
    std::set<si_dataset*> datasets;
    std::set<si_edge_plan*> edge_plans;
};

struct si_buffer {
    si_context* ctx = nullptr;
    void* ptr = nullptr;
    uint64_t bytes = 0;
};

struct si_dataset {
    si_context* ctx = nullptr;
    uint32_t k = 0;
    uint32_t universe = 0;
    uint64_t total = 0;
    uint32_t* d_elems = nullptr;    // total elements (+ 64 padding elements)
    uint32_t* d_sizes = nullptr;    // k
    uint64_t* d_offsets = nullptr;  // k + 1
    std::vector<uint32_t> sizes;    // host copies for planning
    std::vector<uint64_t> offsets;  // k + 1
    uint32_t max_size = 0;
    uint32_t max_elem = 0;          // largest element of the collection
    si_impl::BatchPlan* plan = nullptr;  // batched dispatcher state (built on first use)
    int32_t upload_events[3] = {-1, -1, -1};  // context event ids: transfers, offsets scan, end
    bool validated = false;         // every set checked strictly ascending
};

namespace si_impl {

// Kernel timing bracket: records events around a launch when enabled.
struct KernelScope {
    si_context* ctx;
    const char* name;
    hipStream_t stream;
    hipEvent_t start = nullptr, end = nullptr;
    KernelScope(si_context* c, const char* n);
    KernelScope(si_context* c, const char* n, hipStream_t s);
    ~KernelScope();
};

int check_pair(si_dataset* ds, uint32_t a, uint32_t b);

// Device check that every set of ds is strictly ascending; sets
// ds->validated (SI_ERR_UNSORTED if a set is not).
int validate_dataset(si_dataset* ds);

// Resolves and clears this device's copy of the allocation-free scratch
// array (si_noalloc.hip); called by si_open.
int noalloc_setup(si_context* ctx);

// The technique switch of si_pair_count_async: count of (a, b) into d_out.
int pair_count_dispatch(si_dataset* ds, uint32_t a, uint32_t b, int32_t technique, const si_params* params,
                        uint64_t* d_out);

// Zero-initialised per-context counters region of `bytes` bytes.  Kernels
// that use it leave it zeroed again when they finish.  Inside
// si_pair_count_noalloc it is the scratch array (SI_ERR_RANGE beyond it).
int small_workspace(si_context* ctx, uint64_t bytes, void** out);

// Writes 0 to whichever of the two result slots is given (a pair with an
// empty set).
int zero_outputs(si_context* ctx, uint64_t* d_out64, uint32_t* d_out32);

// ----------------------------------------------------- technique launchers --
// SISI merge path (k_merge.hip)
int ip_pair_count(si_dataset* ds, uint32_t a, uint32_t b, const si_params* p, uint64_t* d_out, uint32_t* d_out32);
int ip_pair_materialize(si_dataset* ds, uint32_t a, uint32_t b, uint32_t* d_out, uint64_t* d_count);
// MISI runs
int ip_run(si_dataset* ds, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts);
// optimized binary search (k_obs.hip)
int obs_pair_count(si_dataset* ds, uint32_t a, uint32_t b, const si_params* p, uint64_t* d_out, uint32_t* d_out32);
// Materialisation by search of the smaller set's elements in the larger one.
int obs_pair_materialize(si_dataset* ds, uint32_t a, uint32_t b, uint32_t* d_out, uint64_t* d_count);
int obs_run(si_dataset* ds, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts);
// hash intersect (k_hash.hip)
int hi_pair_count(si_dataset* ds, uint32_t a, uint32_t b, const si_params* p, uint64_t* d_out, uint32_t* d_out32);
int hi_run(si_dataset* ds, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts);
// --path variants of obs and hi: merge-path segments (k_path.hip); tech is
// SI_TECH_OBS or SI_TECH_HI
int path_pair_count(si_dataset* ds, uint32_t a, uint32_t b, int32_t tech, const si_params* p, uint64_t* d_out,
                    uint32_t* d_out32);
int path_run(si_dataset* ds, int32_t tech, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts);
// bitmaps (k_bitmap.hip)
int bitmap_pair_count(si_dataset* ds, uint32_t a, uint32_t b, const si_params* p, uint64_t* d_out, uint32_t* d_out32);
int bitmap_run(si_dataset* ds, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts);
bool bit_rows_in_lds(uint64_t w64);
int build_bit_rows(si_dataset* ds, uint32_t s0, uint32_t ns, uint64_t w64, unsigned long long* rows, bool zero = true);
int popcount_pairs(si_context* ctx, const unsigned long long* RA, uint32_t ni, const unsigned long long* RB, uint32_t nj,
                   uint64_t w64, bool tri, uint32_t* out, uint64_t ld);
int bitmap_all_pairs(si_dataset* ds, const si_params* p, uint32_t* d_counts);
// bitmap GEMM (k_mm.hip)
int mm_tile_counts(si_dataset* ds, si_tile A, si_tile B, int32_t variant, uint64_t ld, uint32_t* d_C);
// batched dispatcher (k_batched.hip)
int batched_default_tuning(si_dataset* ds, si_tuning* t);
int batched_count(si_dataset* ds, const si_tuning* tuning, uint32_t* d_out);
int batched_autotune(si_dataset* ds, double budget_ms, si_tuning* best);
int batched_autotune_seeded(si_dataset* ds, const si_tuning* seed, double budget_ms, si_tuning* best);
int edge_plan_create(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m, const si_tuning* tuning,
                     si_edge_plan** out);
int edge_count(si_edge_plan* plan, uint32_t* d_counts, uint64_t* d_total);
int edge_count_plain(si_edge_plan* plan, uint32_t* d_counts, uint64_t* d_total);
int edge_autotune(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m, double budget_ms,
                  si_tuning* best);
void edge_plan_free(si_edge_plan* plan);
uint64_t plan_edges(si_edge_plan* plan);
si_dataset* plan_dataset(si_edge_plan* plan);

}  // namespace si_impl
