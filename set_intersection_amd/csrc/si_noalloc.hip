// Allocation-free SISI count (si_pair_count_noalloc).
//
// The count kernels of every SISI technique need device scratch beyond their
// inputs and output: one uint64 partial count per workgroup and a completion
// ticket (the last workgroup folds the partials and resets the ticket to
// zero).  si_pair_count_async takes it from library-allocated memory, grown
// on demand.  This entry point takes it from a file-scope device array of a
// fixed size instead, so it calls no allocator; launches that would need
// more scratch than the array holds are refused.
#include <mutex>

#include "si_internal.h"

namespace {

// Scratch of the allocation-free entry point on every GPU the library runs
// on: bytes [0, 256) hold the kernels' completion ticket (zero between
// launches: the last workgroup of a launch resets it), bytes [256, ...) one
// uint64 partial count per workgroup.  Each GPU holds its own copy until the
// process ends.  In host code the name `scratch` is only a stand-in: the
// device address of a GPU's copy comes from hipGetSymbolAddress.
__device__ __align__(256) unsigned char scratch[SI_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of each GPU's copy

__global__ void k_clear_scratch(uint4* __restrict__ p, uint32_t n16) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Makes the count launchers of this call take their workspace from the
// scratch array and skip the per-kernel timing events.
struct NoAllocScope {
    si_context* ctx;
    explicit NoAllocScope(si_context* c) : ctx(c) { ctx->in_noalloc = true; }
    ~NoAllocScope() { ctx->in_noalloc = false; }
};

}  // namespace

namespace si_impl {

// Called by si_open with the context's device current: resolves this GPU's
// copy of the scratch array once per process and clears it.
int noalloc_setup(si_context* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(SI_ERR_RANGE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[ctx->device]) {
        void* p = nullptr;
        SI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        hipLaunchKernelGGL(k_clear_scratch, dim3(16), dim3(256), 0, ctx->stream, static_cast<uint4*>(p),
                           uint32_t(SI_NOALLOC_SCRATCH_BYTES / 16));
        SI_HIP(hipGetLastError());
        SI_HIP(hipStreamSynchronize(ctx->stream));
        g_scratch[ctx->device] = p;
    }
    ctx->noalloc_ws = g_scratch[ctx->device];
    return SI_OK;
}

}  // namespace si_impl

using namespace si_impl;

extern "C" int si_pair_count_noalloc(si_dataset* ds, uint32_t a, uint32_t b, int32_t technique,
                                     const si_params* params, si_buffer* out, uint64_t offset) {
    SI_TRY(check_pair(ds, a, b));
    if (!out || offset + 8 > out->bytes || (offset & 7)) return fail(SI_ERR_ARG, "bad output buffer/offset");
    si_context* ctx = ds->ctx;
    if (!ctx->noalloc_ws) return fail(SI_ERR_HIP, "allocation-free scratch not set up on this device");
    uint64_t* d = reinterpret_cast<uint64_t*>(static_cast<char*>(out->ptr) + offset);
    NoAllocScope scope(ctx);
    return pair_count_dispatch(ds, a, b, technique, params, d);
}
