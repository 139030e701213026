// libsi.so runtime: contexts, device memory, datasets, event timing, and the
// C-ABI entry points that dispatch to the technique launchers.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "si_device.h"
#include "si_internal.h"

namespace si_impl {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

int DevMem::ensure(uint64_t want) {
    if (want <= bytes && ptr) return SI_OK;
    release();
    if (want == 0) want = 256;
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return fail(SI_ERR_NOMEM, "hipMalloc(" + std::to_string(want) + " B) failed: " + hipGetErrorString(e));
    }
    bytes = want;
    return SI_OK;
}

void DevMem::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

static hipEvent_t take_event(si_context* ctx) {
    if (!ctx->free_events.empty()) {
        hipEvent_t e = ctx->free_events.back();
        ctx->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Folds finished kernel timings into the per-name sums.
static void drain_pending(si_context* ctx) {
    for (auto& p : ctx->pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.end) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.end) == hipSuccess) {
            auto& s = ctx->kstats[p.name];
            s.ms += ms;
            s.launches += 1;
        }
        ctx->free_events.push_back(p.start);
        ctx->free_events.push_back(p.end);
    }
    ctx->pending.clear();
}

KernelScope::KernelScope(si_context* c, const char* n) : KernelScope(c, n, c->stream) {}

KernelScope::KernelScope(si_context* c, const char* n, hipStream_t s) : ctx(c), name(n), stream(s) {
    if (!ctx->kernel_timing || ctx->in_noalloc) return;  // the allocation-free entry point creates no events
    start = take_event(ctx);
    end = take_event(ctx);
    if (start) (void)hipEventRecord(start, stream);
}

KernelScope::~KernelScope() {
    if (!ctx->kernel_timing || !start || !end) return;
    (void)hipEventRecord(end, stream);
    ctx->pending.push_back({name, start, end});
    if (ctx->pending.size() > 4096) drain_pending(ctx);
}

int check_pair(si_dataset* ds, uint32_t a, uint32_t b) {
    if (!ds) return fail(SI_ERR_ARG, "null dataset");
    if (a >= ds->k || b >= ds->k)
        return fail(SI_ERR_ARG, "set index out of range (k=" + std::to_string(ds->k) + ")");
    return SI_OK;
}

int small_workspace(si_context* ctx, uint64_t bytes, void** out) {
    if (ctx->in_noalloc) {
        if (bytes > SI_NOALLOC_SCRATCH_BYTES)
            return fail(SI_ERR_RANGE, "the launch needs " + std::to_string(bytes) +
                                          " B of scratch; the allocation-free entry point has " +
                                          std::to_string(SI_NOALLOC_SCRATCH_BYTES) + " B (lower si_params.blocks)");
        *out = ctx->noalloc_ws;
        return SI_OK;
    }
    if (bytes > ctx->ws_small.bytes || !ctx->ws_small.ptr) {
        uint64_t want = std::max<uint64_t>(bytes, 1 << 16);
        SI_TRY(ctx->ws_small.ensure(want));
        ctx->ws_small_zeroed = false;
    }
    if (!ctx->ws_small_zeroed) {
        SI_HIP(hipMemsetAsync(ctx->ws_small.ptr, 0, ctx->ws_small.bytes, ctx->stream));
        ctx->ws_small_zeroed = true;
    }
    *out = ctx->ws_small.ptr;
    return SI_OK;
}

// (a one-lane kernel rather than hipMemsetAsync: memset nodes have given
// wrong results on graph replays)
__global__ void k_zero_outputs(uint64_t* __restrict__ out64, uint32_t* __restrict__ out32) {
    if (out64) *out64 = 0;
    if (out32) *out32 = 0;
}

int zero_outputs(si_context* ctx, uint64_t* d_out64, uint32_t* d_out32) {
    hipLaunchKernelGGL(k_zero_outputs, dim3(1), dim3(1), 0, ctx->stream, d_out64, d_out32);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

}  // namespace si_impl

using namespace si_impl;

// ------------------------------------------------------------ validation --
namespace {

__global__ void k_offsets_scan(const uint32_t* __restrict__ sizes, uint32_t k, uint64_t* __restrict__ offsets) {
    // single workgroup, chunked exclusive scan with 64-bit carry
    __shared__ uint64_t carry;
    __shared__ uint64_t wsum[16];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t base = 0; base < k; base += blockDim.x) {
        const uint64_t i = base + threadIdx.x;
        const uint32_t v = i < k ? sizes[i] : 0u;
        uint64_t x = v;
        // wave inclusive scan in 64 bits
        for (int off = 1; off < 64; off <<= 1) {
            uint64_t t = __shfl_up(x, off, 64);
            if ((threadIdx.x & 63) >= uint32_t(off)) x += t;
        }
        const int w = threadIdx.x / 64;
        if ((threadIdx.x & 63) == 63) wsum[w] = x;
        __syncthreads();
        uint64_t before = carry;
        for (int j = 0; j < w; ++j) before += wsum[j];
        if (i < k) offsets[i] = before + x - v;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) carry = before + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) offsets[k] = carry;
}

__global__ void k_check_sorted(const uint32_t* __restrict__ elems, const uint64_t* __restrict__ offsets, uint32_t k,
                               uint64_t total, uint32_t* __restrict__ bad) {
    // element i is bad if it is not greater than its predecessor in the same set
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x + 1; i < total; i += stride) {
        if (elems[i] <= elems[i - 1]) {
            // is i the first element of a set?  binary search offsets
            uint32_t lo = 0, hi = k;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (offsets[mid] <= i) lo = mid + 1; else hi = mid;
            }
            // lo-1 is the set containing i; i is its first element iff offsets[lo-1] == i
            if (lo == 0 || offsets[lo - 1] != i) atomicAdd(bad, 1u);
        }
    }
}

__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" {

const char* si_last_error(void) { return g_last_error.c_str(); }

int si_open(int device, si_context** out) {
    if (!out) return fail(SI_ERR_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    SI_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SI_ERR_ARG, "device " + std::to_string(device) + " not present");
    SI_HIP(hipSetDevice(device));
    auto* ctx = new si_context();
    ctx->device = device;
    hipError_t e = hipGetDeviceProperties(&ctx->prop, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->copy_fork, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (ctx->copy) (void)hipStreamDestroy(ctx->copy);
        if (ctx->join) (void)hipEventDestroy(ctx->join);
        if (ctx->fork) (void)hipEventDestroy(ctx->fork);
        if (ctx->side) (void)hipStreamDestroy(ctx->side);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(SI_ERR_HIP, std::string("si_open: ") + hipGetErrorString(e));
    }
    ctx->num_cus = ctx->prop.multiProcessorCount;
    if (noalloc_setup(ctx) != SI_OK) {
        const std::string msg = si_last_error();
        si_close(ctx);
        return fail(SI_ERR_HIP, "si_open: " + msg);
    }
    // tile shape of the SISI merge kernel for materialisation (tuning sweeps)
    if (const char* e = std::getenv("SI_SISI_SHAPE")) ctx->sisi_shape = std::atoi(e);
    if (const char* e = std::getenv("SI_SISI_PRIO")) ctx->sisi_prio = std::atoi(e);
    if (const char* e = std::getenv("SI_SISI_RAMP")) ctx->sisi_ramp = std::atoi(e);
    *out = ctx;
    return SI_OK;
}

int si_close(si_context* ctx) {
    if (!ctx) return SI_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->copy);
    // handles created from the context die with it
    while (!ctx->edge_plans.empty()) si_edge_plan_free(*ctx->edge_plans.begin());
    while (!ctx->datasets.empty()) si_dataset_free(*ctx->datasets.begin());
    while (!ctx->buffers.empty()) si_buffer_free(*ctx->buffers.begin());
    drain_pending(ctx);
    for (auto e : ctx->events) (void)hipEventDestroy(e);
    for (auto e : ctx->free_events) (void)hipEventDestroy(e);
    ctx->ws_small.release();
    ctx->ws_stage.release();
    ctx->ws_lookback.release();
    ctx->ws_aux.release();
    ctx->ws_aux2.release();
    for (void* h : ctx->host_allocs) (void)hipHostFree(h);
    ctx->host_allocs.clear();
    (void)hipStreamSynchronize(ctx->side);
    (void)hipEventDestroy(ctx->copy_fork);
    (void)hipStreamDestroy(ctx->copy);
    (void)hipEventDestroy(ctx->fork);
    (void)hipEventDestroy(ctx->join);
    (void)hipStreamDestroy(ctx->side);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SI_OK;
}

int si_device_info_get(si_context* ctx, si_device_info* out) {
    if (!ctx || !out) return fail(SI_ERR_ARG, "null argument");
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, ctx->prop.gcnArchName, sizeof(out->name) - 1);
    out->multiprocessor_count = ctx->prop.multiProcessorCount;
    out->max_threads_per_block = ctx->prop.maxThreadsPerBlock;
    out->warp_size = ctx->prop.warpSize;
    out->lds_per_block = int32_t(ctx->prop.sharedMemPerBlock);
    out->total_global_mem = ctx->prop.totalGlobalMem;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) out->free_global_mem = fr;
    out->clock_khz = ctx->prop.clockRate;
    out->memory_clock_khz = ctx->prop.memoryClockRate;
    out->memory_bus_width = ctx->prop.memoryBusWidth;
    return SI_OK;
}

int si_sync(si_context* ctx) {
    if (!ctx) return fail(SI_ERR_ARG, "null context");
    SI_HIP(hipStreamSynchronize(ctx->stream));
    SI_HIP(hipStreamSynchronize(ctx->copy));
    return SI_OK;
}

// ---------------------------------------------------------------- events --
int si_event_record(si_context* ctx, int32_t* id) {
    if (!ctx || !id) return fail(SI_ERR_ARG, "null argument");
    hipEvent_t e = nullptr;
    SI_HIP(hipEventCreate(&e));
    SI_HIP(hipEventRecord(e, ctx->stream));
    ctx->events.push_back(e);
    *id = int32_t(ctx->events.size() - 1);
    return SI_OK;
}

int si_event_elapsed(si_context* ctx, int32_t a, int32_t b, float* ms) {
    if (!ctx || !ms) return fail(SI_ERR_ARG, "null argument");
    if (a < 0 || b < 0 || size_t(a) >= ctx->events.size() || size_t(b) >= ctx->events.size())
        return fail(SI_ERR_ARG, "bad event id");
    SI_HIP(hipEventSynchronize(ctx->events[b]));
    SI_HIP(hipEventElapsedTime(ms, ctx->events[a], ctx->events[b]));
    return SI_OK;
}

int si_event_wait(si_context* ctx, int32_t id) {
    if (!ctx) return fail(SI_ERR_ARG, "null context");
    if (id < 0 || size_t(id) >= ctx->events.size()) return fail(SI_ERR_ARG, "bad event id");
    SI_HIP(hipEventSynchronize(ctx->events[id]));
    return SI_OK;
}

int si_event_reset(si_context* ctx) {
    if (!ctx) return fail(SI_ERR_ARG, "null context");
    SI_HIP(hipStreamSynchronize(ctx->stream));
    SI_HIP(hipStreamSynchronize(ctx->copy));
    for (auto e : ctx->events) (void)hipEventDestroy(e);
    ctx->events.clear();
    return SI_OK;
}

int si_kernel_timing(si_context* ctx, int32_t enable) {
    if (!ctx) return fail(SI_ERR_ARG, "null context");
    ctx->kernel_timing = enable != 0;
    return SI_OK;
}

int si_kernel_stats(si_context* ctx, const char* name, double* total_ms, uint64_t* launches) {
    if (!ctx || !name) return fail(SI_ERR_ARG, "null argument");
    drain_pending(ctx);
    auto it = ctx->kstats.find(name);
    if (total_ms) *total_ms = it == ctx->kstats.end() ? 0.0 : it->second.ms;
    if (launches) *launches = it == ctx->kstats.end() ? 0 : it->second.launches;
    return SI_OK;
}

int si_kernel_stats_reset(si_context* ctx) {
    if (!ctx) return fail(SI_ERR_ARG, "null context");
    drain_pending(ctx);
    ctx->kstats.clear();
    return SI_OK;
}

// ---------------------------------------------------------------- buffers --
int si_buffer_alloc(si_context* ctx, uint64_t bytes, si_buffer** out) {
    if (!ctx || !out) return fail(SI_ERR_ARG, "null argument");
    *out = nullptr;
    auto* b = new si_buffer();
    b->ctx = ctx;
    b->bytes = bytes;
    hipError_t e = hipMalloc(&b->ptr, bytes ? bytes : 16);
    if (e != hipSuccess) {
        delete b;
        return fail(SI_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
    }
    ctx->buffers.insert(b);
    *out = b;
    return SI_OK;
}

int si_buffer_free(si_buffer* b) {
    if (!b) return SI_OK;
    b->ctx->buffers.erase(b);
    (void)hipStreamSynchronize(b->ctx->stream);
    (void)hipStreamSynchronize(b->ctx->copy);  // an asynchronous download may still read it
    (void)hipFree(b->ptr);
    delete b;
    return SI_OK;
}

int si_buffer_device_ptr(si_buffer* b, void** p) {
    if (!b || !p) return fail(SI_ERR_ARG, "null argument");
    *p = b->ptr;
    return SI_OK;
}

int si_buffer_download(si_buffer* b, uint64_t off, void* host, uint64_t bytes) {
    if (!b || (!host && bytes)) return fail(SI_ERR_ARG, "null argument");
    if (off + bytes > b->bytes) return fail(SI_ERR_RANGE, "download beyond buffer end");
    if (!bytes) return SI_OK;
    SI_HIP(hipMemcpyAsync(host, static_cast<char*>(b->ptr) + off, bytes, hipMemcpyDeviceToHost, b->ctx->stream));
    SI_HIP(hipStreamSynchronize(b->ctx->stream));
    return SI_OK;
}

int si_buffer_download_async(si_buffer* b, uint64_t off, void* host, uint64_t bytes, int32_t* start_id,
                             int32_t* end_id) {
    if (!b || (!host && bytes)) return fail(SI_ERR_ARG, "null argument");
    if (off + bytes > b->bytes) return fail(SI_ERR_RANGE, "download beyond buffer end");
    si_context* ctx = b->ctx;
    // the copy stream waits for everything enqueued so far on the context stream
    SI_HIP(hipEventRecord(ctx->copy_fork, ctx->stream));
    SI_HIP(hipStreamWaitEvent(ctx->copy, ctx->copy_fork, 0));
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (auto& e : ev) {
        SI_HIP(hipEventCreate(&e));
        ctx->events.push_back(e);  // the context's event table owns it from here on
    }
    SI_HIP(hipEventRecord(ev[0], ctx->copy));
    if (bytes)
        SI_HIP(hipMemcpyAsync(host, static_cast<char*>(b->ptr) + off, bytes, hipMemcpyDeviceToHost, ctx->copy));
    SI_HIP(hipEventRecord(ev[1], ctx->copy));
    if (start_id) *start_id = int32_t(ctx->events.size() - 2);
    if (end_id) *end_id = int32_t(ctx->events.size() - 1);
    return SI_OK;
}

int si_host_alloc(si_context* ctx, uint64_t bytes, void** out) {
    if (!ctx || !out) return fail(SI_ERR_ARG, "null argument");
    *out = nullptr;
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes ? bytes : 16, hipHostMallocDefault);
    if (e != hipSuccess)
        return fail(SI_ERR_NOMEM, "hipHostMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
    ctx->host_allocs.insert(p);
    *out = p;
    return SI_OK;
}

int si_host_free(si_context* ctx, void* p) {
    if (!ctx) return fail(SI_ERR_ARG, "null context");
    if (!p) return SI_OK;
    auto it = ctx->host_allocs.find(p);
    if (it == ctx->host_allocs.end()) return fail(SI_ERR_ARG, "pointer not from si_host_alloc on this context");
    SI_HIP(hipStreamSynchronize(ctx->copy));  // a pending download may still write it
    ctx->host_allocs.erase(it);
    SI_HIP(hipHostFree(p));
    return SI_OK;
}

int si_buffer_upload(si_buffer* b, uint64_t off, const void* host, uint64_t bytes) {
    if (!b || (!host && bytes)) return fail(SI_ERR_ARG, "null argument");
    if (off + bytes > b->bytes) return fail(SI_ERR_RANGE, "upload beyond buffer end");
    if (!bytes) return SI_OK;
    SI_HIP(hipMemcpyAsync(static_cast<char*>(b->ptr) + off, host, bytes, hipMemcpyHostToDevice, b->ctx->stream));
    SI_HIP(hipStreamSynchronize(b->ctx->stream));
    return SI_OK;
}

int si_buffer_zero(si_buffer* b) {
    if (!b) return fail(SI_ERR_ARG, "null buffer");
    SI_HIP(hipMemsetAsync(b->ptr, 0, b->bytes, b->ctx->stream));
    return SI_OK;
}

int si_buffer_copy(si_buffer* dst, const si_buffer* src, uint64_t bytes) {
    if (!dst || !src) return fail(SI_ERR_ARG, "null buffer");
    if ((bytes & 15) || bytes > dst->bytes || bytes > src->bytes) return fail(SI_ERR_RANGE, "bad copy size");
    si_context* ctx = dst->ctx;
    const uint64_t n16 = bytes / 16;
    const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((n16 + 255) / 256, uint64_t(ctx->num_cus) * 32)));
    KernelScope ks(ctx, "buffer_copy");
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, ctx->stream, static_cast<const uint4*>(src->ptr),
                       static_cast<uint4*>(dst->ptr), n16);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

// ---------------------------------------------------------------- dataset --
int si_dataset_upload(si_context* ctx, uint32_t k, uint32_t universe, uint64_t total, const uint32_t* sizes,
                      const uint32_t* elements, int32_t validate, si_dataset** out) {
    if (!ctx || !out || (k && !sizes) || (total && !elements)) return fail(SI_ERR_ARG, "null argument");
    *out = nullptr;
    uint64_t sum = 0;
    uint32_t mx = 0;
    for (uint32_t i = 0; i < k; ++i) {
        sum += sizes[i];
        mx = std::max(mx, sizes[i]);
    }
    if (sum != total)
        return fail(SI_ERR_ARG, "sum of sizes (" + std::to_string(sum) + ") != total (" + std::to_string(total) + ")");
    auto* ds = new si_dataset();
    ds->ctx = ctx;
    ds->k = k;
    ds->universe = universe;
    ds->total = total;
    ds->max_size = mx;
    ds->sizes.assign(sizes, sizes + k);
    ds->offsets.assign(size_t(k) + 1, 0);
    for (uint32_t i = 0; i < k; ++i) ds->offsets[i + 1] = ds->offsets[i] + sizes[i];
    auto cleanup = [&](int st, const std::string& m) {
        si_dataset_free(ds);
        return fail(st, m);
    };
    // 64 padding elements: 16-byte chunk loads may run up to 3 elements past the end
    const uint64_t elemBytes = (total + 64) * sizeof(uint32_t);
    if (hipMalloc(&ds->d_elems, elemBytes) != hipSuccess) return cleanup(SI_ERR_NOMEM, "hipMalloc(elements) failed");
    if (hipMalloc(&ds->d_sizes, (uint64_t(k) + 1) * 4) != hipSuccess) return cleanup(SI_ERR_NOMEM, "hipMalloc(sizes) failed");
    if (hipMalloc(&ds->d_offsets, (uint64_t(k) + 1) * 8) != hipSuccess)
        return cleanup(SI_ERR_NOMEM, "hipMalloc(offsets) failed");
    // events bracketing the transfers and the offsets scan (si_dataset_upload_events)
    if (si_event_record(ctx, &ds->upload_events[0]) != SI_OK) return cleanup(SI_ERR_HIP, si_last_error());
    hipError_t e = hipMemsetAsync(ds->d_elems + total, 0, 64 * 4, ctx->stream);
    if (e == hipSuccess && total)
        e = hipMemcpyAsync(ds->d_elems, elements, total * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && k) e = hipMemcpyAsync(ds->d_sizes, sizes, uint64_t(k) * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return cleanup(SI_ERR_HIP, std::string("dataset upload: ") + hipGetErrorString(e));
    if (si_event_record(ctx, &ds->upload_events[1]) != SI_OK) return cleanup(SI_ERR_HIP, si_last_error());
    {
        KernelScope ks(ctx, "offsets_scan");
        hipLaunchKernelGGL(k_offsets_scan, dim3(1), dim3(1024), 0, ctx->stream, ds->d_sizes, k, ds->d_offsets);
    }
    e = hipGetLastError();
    if (e == hipSuccess && si_event_record(ctx, &ds->upload_events[2]) != SI_OK) return cleanup(SI_ERR_HIP, si_last_error());
    if (e != hipSuccess) return cleanup(SI_ERR_HIP, std::string("offsets scan: ") + hipGetErrorString(e));
    // the largest element (sets are sorted: the max of the last elements)
    for (uint32_t i = 0; i < k; ++i)
        if (sizes[i]) ds->max_elem = std::max(ds->max_elem, elements[ds->offsets[i + 1] - 1]);
    if (validate) {
        const int st = validate_dataset(ds);
        if (st != SI_OK) return cleanup(st, si_last_error());
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return cleanup(SI_ERR_HIP, std::string("dataset upload: ") + hipGetErrorString(e));
    ctx->datasets.insert(ds);
    *out = ds;
    return SI_OK;
}

}  // extern "C"

namespace si_impl {

// Checks on the device that every set is strictly ascending
// (SI_ERR_UNSORTED otherwise) and marks the dataset validated.
int validate_dataset(si_dataset* ds) {
    si_context* ctx = ds->ctx;
    if (ds->total > 1) {
        void* ws = nullptr;
        SI_TRY(small_workspace(ctx, 256, &ws));
        uint32_t* bad = static_cast<uint32_t*>(ws) + 32;  // slot reserved for validation
        const uint32_t grid = uint32_t(std::min<uint64_t>((ds->total + 255) / 256, uint64_t(ctx->num_cus) * 16));
        hipLaunchKernelGGL(k_check_sorted, dim3(grid), dim3(256), 0, ctx->stream, ds->d_elems, ds->d_offsets, ds->k,
                           ds->total, bad);
        uint32_t nbad = 0;
        SI_HIP(hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, ctx->stream));
        SI_HIP(hipMemsetAsync(bad, 0, 4, ctx->stream));
        SI_HIP(hipStreamSynchronize(ctx->stream));
        if (nbad)
            return fail(SI_ERR_UNSORTED, std::to_string(nbad) + " elements are not strictly ascending within their set");
    }
    ds->validated = true;
    return SI_OK;
}

}  // namespace si_impl

extern "C" {

int si_dataset_free(si_dataset* ds) {
    if (!ds) return SI_OK;
    ds->ctx->datasets.erase(ds);
    // edge plans read the dataset: they go first
    std::vector<si_edge_plan*> dependent;
    for (si_edge_plan* p : ds->ctx->edge_plans)
        if (si_impl::plan_dataset(p) == ds) dependent.push_back(p);
    for (si_edge_plan* p : dependent) si_edge_plan_free(p);
    (void)hipStreamSynchronize(ds->ctx->stream);
    if (ds->plan) free_batch_plan(ds->plan);
    if (ds->d_elems) (void)hipFree(ds->d_elems);
    if (ds->d_sizes) (void)hipFree(ds->d_sizes);
    if (ds->d_offsets) (void)hipFree(ds->d_offsets);
    delete ds;
    return SI_OK;
}

int si_dataset_upload_events(si_dataset* ds, int32_t* transfer_start, int32_t* scan_start, int32_t* scan_end) {
    if (!ds) return fail(SI_ERR_ARG, "null dataset");
    if (transfer_start) *transfer_start = ds->upload_events[0];
    if (scan_start) *scan_start = ds->upload_events[1];
    if (scan_end) *scan_end = ds->upload_events[2];
    return SI_OK;
}

int si_dataset_info(si_dataset* ds, uint32_t* k, uint32_t* universe, uint64_t* total) {
    if (!ds) return fail(SI_ERR_ARG, "null dataset");
    if (k) *k = ds->k;
    if (universe) *universe = ds->universe;
    if (total) *total = ds->total;
    return SI_OK;
}

}  // extern "C"
