/*
 * set_intersection.h -- C ABI of the MI355X (gfx950) set-intersection library
 * (libsi.so).
 *
 * The reference project (chribell/set_intersection) has no library API: its
 * boundary is one executable per technique that reads the .bin dataset format
 * (include/io.hpp:25-54), runs CUDA kernels, and writes "(i,j): count" lines
 * (include/io.hpp:91-136).  This header is the thin C layer those executables
 * call instead of the CUDA runtime: every entry point below replaces a block of
 * CUDA host code in one of the reference's main() functions, cited per
 * function.  The per-technique executables under set_intersection_amd/apps
 * call only this header.
 *
 * ABI rules
 *   - Plain C: fixed-width integers, opaque handles, no exceptions cross it.
 *   - Every function returns SI_OK (0) or a negative si_status; the message of
 *     the last failure on the calling thread is returned by si_last_error().
 *   - Host buffers are owned by the caller.  Device memory is owned by the
 *     library (si_dataset, si_buffer) and released by the matching free call.
 *   - One host thread per si_context.  All work of a context is issued on that
 *     context's single HIP stream, in call order.
 *   - Functions whose name ends in _async only enqueue work and write their
 *     results into device buffers (si_buffer); everything else synchronises
 *     the stream before returning.
 *   - Elements are uint32 and must be sorted ascending and unique inside each
 *     set (the reference's generators and txt2bin produce such sets).  The
 *     full uint32 range [0, 2^32) is supported; all index arithmetic is 64-bit.
 */
#ifndef SET_INTERSECTION_H
#define SET_INTERSECTION_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status -- */
typedef enum si_status {
    SI_OK = 0,
    SI_ERR_ARG = -1,        /* invalid argument (null pointer, bad index, ...) */
    SI_ERR_HIP = -2,        /* a HIP runtime call failed                        */
    SI_ERR_NOMEM = -3,      /* device or host allocation failed                 */
    SI_ERR_RANGE = -4,      /* a size exceeds what the entry point supports     */
    SI_ERR_UNSORTED = -5,   /* a set is not strictly ascending                  */
    SI_ERR_CAPACITY = -6,   /* an output buffer given by the caller is too small*/
    SI_ERR_DEVICE = -7      /* a kernel reported an internal error              */
} si_status;

/* Message of the last failed call on this thread ("" if none). */
const char* si_last_error(void);

/* ------------------------------------------------------------ techniques -- */
/* One value per GPU executable of the reference (CMakeLists.txt:99-152). */
typedef enum si_technique {
    SI_TECH_IP = 0,             /* src/ip.cu             merge path             */
    SI_TECH_OBS = 1,            /* src/obs.cu            optimized binary search*/
    SI_TECH_HI = 2,             /* src/hi.cu             hash intersect         */
    SI_TECH_BITMAP_NAIVE = 3,   /* src/bitmap_naive.cu   bitmap AND + popcount  */
    SI_TECH_BITMAP_DYNAMIC = 4, /* src/bitmap_dynamic.cu per-set bitmap probe   */
    SI_TECH_MM = 5,             /* src/matrix_multiplication.cu bitmap GEMM     */
    SI_TECH_COUNT_ = 6
} si_technique;

/* Launch parameters: the reference's CLI flags (ip.cu:31-43, obs.cu:55-70,
 * hi.cu:67-86, bitmap_dynamic.cu:41-51).  Zero means "library default". */
typedef struct si_params {
    uint32_t blocks;        /* --blocks   (reference default: 16 x CUs)         */
    uint32_t threads;       /* --threads  (reference default: maxThreads/2)     */
    uint32_t warp;          /* --warp     obs/hi: wave-level instead of block   */
    uint32_t path;          /* --path     obs/hi: cut the merge path of each pair
                             * into segments and apply the technique per
                             * segment (intersectPathOBS obs.cu:438-522,
                             * intersectPathHI hi.cu:420-479); SISI pairs and
                             * runs; ignored by ip and bitmap_dynamic       */
    uint32_t buckets;       /* --buckets      hi (default 512)                  */
    uint32_t bucket_size;   /* --bucket-size  hi (default 1000)                 */
    uint32_t reserved[4];   /* tuning sweeps, 0 = automatic: [0] ip SISI
                             * tile shape; [1] bitmap_dynamic SISI bucket
                             * width (log2 values per bucket, 10..18)      */
} si_params;

/* A contiguous range of set ids [start, end) with tile number id; mirrors
 * struct tile (include/util.hpp:11-20). */
typedef struct si_tile {
    uint32_t id;
    uint32_t start;
    uint32_t end;
} si_tile;

/* --------------------------------------------------------------- context -- */
typedef struct si_context si_context;

typedef struct si_device_info {
    char name[64];              /* gcnArchName, e.g. "gfx950:sramecc+:xnack-"   */
    int32_t multiprocessor_count;
    int32_t max_threads_per_block;
    int32_t warp_size;          /* 64 on gfx950 */
    int32_t lds_per_block;      /* bytes */
    uint64_t total_global_mem;
    uint64_t free_global_mem;
    int32_t clock_khz;
    int32_t memory_clock_khz;
    int32_t memory_bus_width;
} si_device_info;

/* Opens `device` and creates the context's stream.  Replaces the
 * cudaDeviceGetAttribute calls at ip.cu:25-26 (and the same lines of every
 * .cu main). */
int si_open(int device, si_context** out);
int si_close(si_context* ctx);
int si_device_info_get(si_context* ctx, si_device_info* out);
/* Waits until all work enqueued on the context has finished. */
int si_sync(si_context* ctx);

/* --------------------------------------------------------- device events -- */
/* Named timing of stream work, the device side of DeviceTimer
 * (include/device_timer.cuh:12-77).  si_event_record enqueues a HIP event on
 * the context stream and returns its id; si_event_elapsed returns the time in
 * milliseconds between two recorded events (synchronising on the later one). */
int si_event_record(si_context* ctx, int32_t* event_id);
int si_event_elapsed(si_context* ctx, int32_t start_id, int32_t end_id, float* ms);
int si_event_reset(si_context* ctx);

/* Per-kernel timing: when enabled, every kernel the library launches is
 * bracketed by HIP events on the context stream and its duration is summed
 * under the kernel's name.  si_kernel_stats copies the sum (ms) and launch
 * count for `name` (0 launches if the name never ran). */
int si_kernel_timing(si_context* ctx, int32_t enable);
int si_kernel_stats(si_context* ctx, const char* name, double* total_ms, uint64_t* launches);
int si_kernel_stats_reset(si_context* ctx);

/* ---------------------------------------------------------- device memory -- */
typedef struct si_buffer si_buffer;
int si_buffer_alloc(si_context* ctx, uint64_t bytes, si_buffer** out);
int si_buffer_free(si_buffer* buf);
/* Raw device address (for handing to another HIP component). */
int si_buffer_device_ptr(si_buffer* buf, void** dptr);
int si_buffer_download(si_buffer* buf, uint64_t offset, void* host, uint64_t bytes);
int si_buffer_upload(si_buffer* buf, uint64_t offset, const void* host, uint64_t bytes);
int si_buffer_zero(si_buffer* buf);
/* Device-to-device copy of `bytes` (multiple of 16) by a plain streaming
 * kernel ("buffer_copy"); its HBM counters calibrate traffic measurements. */
int si_buffer_copy(si_buffer* dst, const si_buffer* src, uint64_t bytes);

/* ---------------------------------------------------- pipelined transfers -- */
/* The reference copies each run's counts to the host with a blocking
 * cudaMemcpy between the run's kernels and the next run's (ip.cu:166-169;
 * the same step in obs.cu, hi.cu, bitmap_dynamic.cu and
 * matrix_multiplication.cu), so the GPU idles during every transfer.  These
 * calls let a driver overlap run i's transfer (and the host's writing of run
 * i's result lines) with run i+1's kernels.
 *
 * si_host_alloc: page-locked host memory (released by si_host_free or at
 * si_close). */
int si_host_alloc(si_context* ctx, uint64_t bytes, void** out);
int si_host_free(si_context* ctx, void* host);
/* Enqueues a device-to-host copy of `bytes` at `offset` of `buf` on the
 * context's copy stream, ordered after all work enqueued so far on the
 * context stream, and returns without waiting.  *start_event / *end_event
 * (either may be null) receive ids of events bracketing the copy, usable with
 * si_event_elapsed; the host bytes are final once si_event_wait(end) returns.
 * Work enqueued later on the context stream does not wait for the copy: do not
 * write to `buf` again (e.g. by launching another run into it) before the copy
 * has ended.  `host` should come from si_host_alloc; pageable memory is copied
 * correctly but without overlap. */
int si_buffer_download_async(si_buffer* buf, uint64_t offset, void* host, uint64_t bytes,
                             int32_t* start_event, int32_t* end_event);
/* Blocks the calling thread until the event (si_event_record,
 * si_buffer_download_async) has completed. */
int si_event_wait(si_context* ctx, int32_t event_id);

/* ---------------------------------------------------------------- dataset -- */
typedef struct si_dataset si_dataset;

/* Copies a collection to HBM: sizes[k], elements[total] exactly as in the .bin
 * file (include/io.hpp:25-54) and computes k+1 64-bit offsets on the device.
 * Replaces the H2D copies and thrust::exclusive_scan at ip.cu:120-137 (same in
 * obs.cu:152-171, hi.cu:175-196, bitmap_naive.cu:93-110,
 * bitmap_dynamic.cu:328-361).  If `validate` is non-zero every set is checked
 * to be strictly ascending (SI_ERR_UNSORTED otherwise). */
int si_dataset_upload(si_context* ctx, uint32_t k, uint32_t universe, uint64_t total,
                      const uint32_t* sizes, const uint32_t* elements, int32_t validate,
                      si_dataset** out);
int si_dataset_free(si_dataset* ds);
/* Ids of the context events si_dataset_upload recorded on the context stream:
 * before the host-to-device copies, between the copies and the offsets scan,
 * and after the scan.  With si_event_elapsed they time the reference's
 * "Transfer to device" and "Compute set offsets" rows separately
 * (ip.cu:128-137; obs.cu:162-171).  The ids are invalid after si_event_reset. */
int si_dataset_upload_events(si_dataset* ds, int32_t* transfer_start, int32_t* scan_start, int32_t* scan_end);
int si_dataset_info(si_dataset* ds, uint32_t* k, uint32_t* universe, uint64_t* total);

/* ------------------------------------------------------------------ SISI -- */
/* |S_a ∩ S_b| for one pair of sets.  technique is SI_TECH_IP (merge path:
 * findDiagonals + intersectPath, helpers.cuh:74-166 and ip.cu:207-245),
 * SI_TECH_OBS (obs.cu:358-436), SI_TECH_HI (hi.cu:361-418) or
 * SI_TECH_BITMAP_DYNAMIC (constructBitmap + singleInstanceIntersect,
 * bitmap_dynamic.cu:240-276).  The _async form writes the count as one uint64
 * at byte offset `offset` of `out`. */
int si_pair_count(si_dataset* ds, uint32_t a, uint32_t b, int32_t technique,
                  const si_params* params, uint64_t* count);
int si_pair_count_async(si_dataset* ds, uint32_t a, uint32_t b, int32_t technique,
                        const si_params* params, si_buffer* out, uint64_t offset);

/* Allocation-free form of si_pair_count_async: same arguments, techniques
 * and result.  No counterpart in the reference; it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The count kernels' scratch (one uint64
 * per workgroup plus a completion ticket; for obs also the searched set's
 * 16 KiB of gathered samples) comes from a file-scope device array of
 * SI_NOALLOC_SCRATCH_BYTES per GPU, set up by si_open, instead of
 * library-allocated memory; no allocator is called.  A launch that would need
 * more scratch (more than 8160 workgroups, 6112 for obs, e.g. from a large
 * si_params.blocks) fails with SI_ERR_RANGE.  All calls on one GPU share the
 * array: calls that overlap in time on one GPU (other contexts, other host
 * threads) race and must be serialised by the caller. */
#define SI_NOALLOC_SCRATCH_BYTES (64u * 1024u)
int si_pair_count_noalloc(si_dataset* ds, uint32_t a, uint32_t b, int32_t technique,
                          const si_params* params, si_buffer* out, uint64_t offset);

/* Materialises S_a ∩ S_b in ascending order (the std::set_intersection output
 * vector of std_set_intersection.cpp:68-72).  `cap` is the capacity of `out`
 * in elements; *n receives the intersection size.  SI_ERR_CAPACITY (with *n
 * set) if cap < *n. */
int si_pair_materialize(si_dataset* ds, uint32_t a, uint32_t b, uint32_t* out,
                        uint64_t cap, uint64_t* n);
/* Device form: writes the list to `out` (device) and the size as uint64 at
 * byte offset 0 of `count_out`.  `out` must hold min(|S_a|,|S_b|) elements.
 * Sets of similar size are merged (the merge-path kernel of si_pair_count
 * with SI_TECH_IP); when one set is at least 64 times the other, the smaller
 * set's elements are searched in the larger one instead (the obs search of
 * obs.cu:296-345, with a decoupled look-back placing the matches in order).
 * A dataset uploaded with validate = 0 is validated (once, synchronously) by
 * its first materialisation: SI_ERR_UNSORTED if a set is not strictly
 * ascending. */
int si_pair_materialize_async(si_dataset* ds, uint32_t a, uint32_t b, si_buffer* out,
                              si_buffer* count_out);

/* ------------------------------------------------------------------ MISI -- */
/* One GPU invocation ("run") of the reference's multi-run driver
 * (ip.cu:141-177 and the same loop in obs.cu, hi.cu, bitmap_dynamic.cu):
 * counts for all pairs (i in A, j in B, j > i when A.id == B.id), written in
 * the reference layout: counts[triangular_index(n, i - A.id*n, j - B.id*n)]
 * for a self-join, counts[quadratic_index(n, ...)] otherwise, with
 * n = num_sets (include/util.hpp:68-74).  counts_host must hold
 * n*(n-1)/2 (self-join) or n*n entries; untouched slots are zeroed.
 * technique: SI_TECH_IP, SI_TECH_OBS, SI_TECH_HI, SI_TECH_BITMAP_DYNAMIC. */
int si_run_count(si_dataset* ds, int32_t technique, const si_params* params,
                 si_tile A, si_tile B, uint32_t num_sets, uint32_t* counts_host);
int si_run_count_async(si_dataset* ds, int32_t technique, const si_params* params,
                       si_tile A, si_tile B, uint32_t num_sets, si_buffer* counts);

/* bitmap_naive (bitmap_naive.cu:93-126): bitmaps of all k sets, then AND +
 * popcount for all k(k-1)/2 pairs, counts[triangular_index(k, a, b)]. */
int si_all_pairs_bitmap(si_dataset* ds, const si_params* params, uint32_t* counts_host);
int si_all_pairs_bitmap_async(si_dataset* ds, const si_params* params, si_buffer* counts);

/* ------------------------------------------------------------ bitmap GEMM -- */
/* The mm executable (matrix_multiplication.cu:139-184): counts of tile pair
 * (A, B) as a dense A.len x B.len row-major matrix C[i][j] = |S_{A.start+i} ∩
 * S_{B.start+j}|, computed by an int8 MFMA GEMM of 0/1 incidence matrices
 * built on the device (int32 accumulation: exact below 2^31).
 * variant 0 = int8 MFMA / int32 accumulate, 1 = FP32-accumulate comparison
 * variant (fp8 MFMA), 2 = bitmap AND + popcount, 3 = fp4 (e2m1) MFMA / FP32
 * accumulate, twice the int8 rate.  Variants 1 and 3 are exact while every
 * count is below 2^24 and fail with SI_ERR_RANGE for a universe above 2^24
 * (the reference's cublasSgemm has the same bound).  ld = row stride of C in
 * elements (>= B.len). */
enum { SI_MM_I8 = 0, SI_MM_F32 = 1, SI_MM_POPCOUNT = 2, SI_MM_FP4 = 3 };
int si_mm_counts(si_dataset* ds, si_tile A, si_tile B, int32_t variant, uint32_t ld,
                 uint32_t* counts_host);
int si_mm_counts_async(si_dataset* ds, si_tile A, si_tile B, int32_t variant, uint32_t ld,
                       si_buffer* counts);

/* -------------------------------------------------------- batched MISI -- */
/* All k(k-1)/2 pair counts of the whole collection in triangular order
 * (counts[triangular_index(k, a, b)], the single-run layout of
 * include/io.hpp:91-102).  Has no single reference counterpart: it replaces
 * the per-run loops of ip/obs/hi/bitmap_dynamic (ip.cu:141-177 and the same
 * loop in obs.cu, hi.cu, bitmap_dynamic.cu) for a whole collection.
 *
 * Sets are ordered by size (descending) and every pair is charged to its
 * larger set ("row").  Rows are binned by size and size ratio, and each bin
 * is dispatched to one technique by persistent work-queue kernels:
 *   BITMAP  rows with |S| >= bitmap_min whose universe fits an LDS bitmap;
 *   HASH    rows whose LDS hash table fits (2|S| slots);
 *   SEARCH  larger rows: partners with |row|/|partner| >= ratio_search
 *           binary-search the row (LDS sample cache + global search);
 *   MERGE   the remaining pairs of large rows (merge path).
 * Each work item is (row, range of smaller partner sets); the partner
 * elements are streamed once per row and probed, and per-pair counts are
 * reduced with wave ballots.  `tuning` may be NULL (defaults) or a record from
 * si_batched_autotune. */
typedef struct si_tuning {
    uint32_t version;            /* must be SI_TUNING_VERSION                     */
    uint32_t bitmap_min;         /* rows with |S| >= bitmap_min use LDS bitmaps   */
    uint32_t ratio_search;       /* SEARCH when |row| >= ratio_search * |partner| */
    uint32_t chunk_elems;        /* partner elements per work item (hash rows)    */
    uint32_t chunk_elems_bitmap; /* partner elements per work item (bitmap rows)  */
    uint32_t probe_wg_per_cu;    /* persistent workgroups per CU (hash rows)      */
    uint32_t use_graph;          /* replay the launches from a captured hipGraph  */
    uint32_t edge_bitmap_halves; /* edge plans, universes above 2^19: 1 = bitmap
                                    rows in universe halves (64 KiB bitmaps, two
                                    workgroups per CU), 0 = one 128 KiB bitmap;
                                    universes up to 2^19: 1 = 64 KiB bitmaps,
                                    two workgroups per CU, 0 = one per CU      */
    uint32_t reserved[8];
} si_tuning;
#define SI_TUNING_VERSION 2u
int si_batched_default_tuning(si_dataset* ds, si_tuning* out);
int si_batched_count(si_dataset* ds, const si_tuning* tuning, uint32_t* counts_host);
int si_batched_count_async(si_dataset* ds, const si_tuning* tuning, si_buffer* counts);
/* Times candidate settings on this dataset (each candidate run once, until
 * budget_ms of measured time is spent) and returns the fastest record. */
int si_batched_autotune(si_dataset* ds, double budget_ms, si_tuning* out);
/* As si_batched_autotune, starting the search from `seed` (a record stored
 * for a dataset of similar shape, e.g. the nearest one by size and
 * element-degree statistics) with candidate values around the seed's
 * (seed / 4 .. seed * 4 per knob) instead of the fixed lists.
 * SI_ERR_ARG if the seed is not a valid record. */
int si_batched_autotune_seeded(si_dataset* ds, const si_tuning* seed, double budget_ms, si_tuning* out);

/* ------------------------------------------------------ edge-wise counts -- */
/* For every edge e = (u[e], v[e]) of a graph whose adjacency lists are the
 * sets of `ds` (one set per vertex), counts[e] = |N(u) ∩ N(v)|, plus the
 * uint64 sum over all edges (3x the triangle count when every undirected
 * edge is listed once).  No reference counterpart: the reference intersects
 * all pairs of a collection; an edge list is an explicit pair list handled by
 * the same binned dispatcher (each edge is charged to its endpoint with the
 * larger list, whose bitmap / hash table / sample cache is built once in LDS
 * and probed by the smaller lists of all its edges).
 * si_edge_plan groups the edges (host arrays, O(m)); the plan can be run any
 * number of times. */
typedef struct si_edge_plan si_edge_plan;
int si_edge_plan_create(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m,
                        const si_tuning* tuning, si_edge_plan** out);
int si_edge_plan_free(si_edge_plan* plan);
/* counts: m uint32 (device); total: uint64 at byte offset 0 of total_out. */
int si_edge_count_async(si_edge_plan* plan, si_buffer* counts, si_buffer* total_out);
/* Times candidate tunings of the edge dispatcher on this edge list (one pass
 * of coordinate descent over bitmap_min, chunk sizes, workgroups per CU and
 * the search ratio; each candidate planned, run once to warm up and twice
 * timed, the faster run counting) until budget_ms of wall time is spent, and
 * returns the fastest record for si_edge_plan_create.  With tuning == NULL,
 * si_edge_plan_create uses the edge defaults: bitmap rows from 512
 * elements, 2 Mi partner elements per bitmap-row item, 16384 per hash-row
 * item and two probe workgroups per CU (the other fields as in
 * si_batched_default_tuning). */
int si_edge_autotune(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m, double budget_ms,
                     si_tuning* out);
/* The plain baseline of si_edge_count_async on the same plan and outputs:
 * one wavefront per edge, no grouping by vertex, no LDS tables; each element
 * of the shorter list is binary-searched in the longer list in global memory,
 * the per-pair search of the reference's obs (obs.cu:330-345) applied to an
 * explicit pair list. */
int si_edge_count_plain_async(si_edge_plan* plan, si_buffer* counts, si_buffer* total_out);
/* One-shot form: plan, run, download, free. */
int si_edge_count(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m,
                  const si_tuning* tuning, uint32_t* counts_host, uint64_t* total);

#ifdef __cplusplus
}
#endif

#endif /* SET_INTERSECTION_H */
