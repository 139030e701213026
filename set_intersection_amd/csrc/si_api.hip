// C-ABI entry points for counting: argument checks and technique dispatch.
#include <algorithm>

#include "si_internal.h"

using namespace si_impl;

namespace {

int not_available(const char* what) { return fail(SI_ERR_ARG, std::string(what) + ": technique not supported by this entry point"); }

int check_tile(si_dataset* ds, const si_tile& t) {
    if (t.start > t.end || t.end > ds->k) return fail(SI_ERR_ARG, "tile out of range");
    return SI_OK;
}

uint64_t run_slots(si_tile A, si_tile B, uint32_t n) {
    return (A.id == B.id) ? uint64_t(n) * (n > 0 ? n - 1 : 0) / 2 : uint64_t(n) * n;
}

int check_run(si_dataset* ds, si_tile A, si_tile B, uint32_t n) {
    SI_TRY(check_tile(ds, A));
    SI_TRY(check_tile(ds, B));
    // every local index must fit the n x n layout
    auto fits = [&](const si_tile& t) {
        return t.start >= uint64_t(t.id) * n && (t.end == t.start || uint64_t(t.end) - uint64_t(t.id) * n <= n);
    };
    if (!fits(A) || !fits(B)) return fail(SI_ERR_ARG, "tile does not fit num_sets (local index >= num_sets)");
    if (A.id == B.id && (A.start != B.start || A.end != B.end)) return fail(SI_ERR_ARG, "self-join tiles must be equal");
    return SI_OK;
}

}  // namespace

extern "C" {

int si_pair_count_async(si_dataset* ds, uint32_t a, uint32_t b, int32_t technique, const si_params* params,
                        si_buffer* out, uint64_t offset) {
    SI_TRY(check_pair(ds, a, b));
    if (!out || offset + 8 > out->bytes || (offset & 7)) return fail(SI_ERR_ARG, "bad output buffer/offset");
    uint64_t* d = reinterpret_cast<uint64_t*>(static_cast<char*>(out->ptr) + offset);
    switch (technique) {
        case SI_TECH_IP: return ip_pair_count(ds, a, b, params, d);
        default: return not_available("si_pair_count");
    }
}

int si_pair_count(si_dataset* ds, uint32_t a, uint32_t b, int32_t technique, const si_params* params, uint64_t* count) {
    SI_TRY(check_pair(ds, a, b));
    if (!count) return fail(SI_ERR_ARG, "null count");
    si_context* ctx = ds->ctx;
    SI_TRY(ctx->ws_aux2.ensure(std::max<uint64_t>(ctx->ws_aux2.bytes, 64)));
    si_buffer tmp;
    tmp.ctx = ctx;
    tmp.ptr = ctx->ws_aux2.ptr;
    tmp.bytes = ctx->ws_aux2.bytes;
    SI_TRY(si_pair_count_async(ds, a, b, technique, params, &tmp, 0));
    SI_HIP(hipMemcpyAsync(count, tmp.ptr, 8, hipMemcpyDeviceToHost, ctx->stream));
    SI_HIP(hipStreamSynchronize(ctx->stream));
    return SI_OK;
}

int si_pair_materialize_async(si_dataset* ds, uint32_t a, uint32_t b, si_buffer* out, si_buffer* count_out) {
    SI_TRY(check_pair(ds, a, b));
    if (!out || !count_out || count_out->bytes < 8) return fail(SI_ERR_ARG, "bad output buffers");
    const uint64_t need = std::min(ds->sizes[a], ds->sizes[b]);
    if (out->bytes < need * 4) return fail(SI_ERR_CAPACITY, "output buffer smaller than min(|a|,|b|) elements");
    return ip_pair_materialize(ds, a, b, static_cast<uint32_t*>(out->ptr), static_cast<uint64_t*>(count_out->ptr));
}

int si_pair_materialize(si_dataset* ds, uint32_t a, uint32_t b, uint32_t* out, uint64_t cap, uint64_t* n) {
    SI_TRY(check_pair(ds, a, b));
    if (!n || (!out && cap)) return fail(SI_ERR_ARG, "null argument");
    si_context* ctx = ds->ctx;
    const uint64_t need = std::min(ds->sizes[a], ds->sizes[b]);
    si_buffer* dout = nullptr;
    si_buffer* dcnt = nullptr;
    SI_TRY(si_buffer_alloc(ctx, std::max<uint64_t>(need, 1) * 4, &dout));
    int st = si_buffer_alloc(ctx, 8, &dcnt);
    if (st == SI_OK) st = si_pair_materialize_async(ds, a, b, dout, dcnt);
    uint64_t cnt = 0;
    if (st == SI_OK) st = si_buffer_download(dcnt, 0, &cnt, 8);
    if (st == SI_OK) {
        *n = cnt;
        if (cnt > cap)
            st = fail(SI_ERR_CAPACITY, "intersection has " + std::to_string(cnt) + " elements, capacity " + std::to_string(cap));
        else
            st = si_buffer_download(dout, 0, out, cnt * 4);
    }
    si_buffer_free(dout);
    si_buffer_free(dcnt);
    return st;
}

int si_run_count_async(si_dataset* ds, int32_t technique, const si_params* params, si_tile A, si_tile B,
                       uint32_t num_sets, si_buffer* counts) {
    if (!ds || !counts) return fail(SI_ERR_ARG, "null argument");
    SI_TRY(check_run(ds, A, B, num_sets));
    const uint64_t slots = run_slots(A, B, num_sets);
    if (counts->bytes < slots * 4) return fail(SI_ERR_CAPACITY, "counts buffer too small for the run layout");
    si_context* ctx = ds->ctx;
    SI_HIP(hipMemsetAsync(counts->ptr, 0, slots * 4, ctx->stream));
    uint32_t* d = static_cast<uint32_t*>(counts->ptr);
    switch (technique) {
        case SI_TECH_IP: return ip_run(ds, params, A, B, num_sets, d);
        default: return not_available("si_run_count");
    }
}

int si_run_count(si_dataset* ds, int32_t technique, const si_params* params, si_tile A, si_tile B, uint32_t num_sets,
                 uint32_t* counts_host) {
    if (!ds || !counts_host) return fail(SI_ERR_ARG, "null argument");
    SI_TRY(check_run(ds, A, B, num_sets));
    const uint64_t slots = run_slots(A, B, num_sets);
    si_buffer* d = nullptr;
    SI_TRY(si_buffer_alloc(ds->ctx, std::max<uint64_t>(slots, 1) * 4, &d));
    int st = si_run_count_async(ds, technique, params, A, B, num_sets, d);
    if (st == SI_OK) st = si_buffer_download(d, 0, counts_host, slots * 4);
    si_buffer_free(d);
    return st;
}

// ---- entry points completed in later files (declared in the public ABI) ----
int si_all_pairs_bitmap_async(si_dataset* ds, const si_params*, si_buffer*) {
    (void)ds;
    return not_available("si_all_pairs_bitmap");
}
int si_all_pairs_bitmap(si_dataset* ds, const si_params*, uint32_t*) {
    (void)ds;
    return not_available("si_all_pairs_bitmap");
}
int si_mm_counts_async(si_dataset*, si_tile, si_tile, int32_t, uint32_t, si_buffer*) { return not_available("si_mm_counts"); }
int si_mm_counts(si_dataset*, si_tile, si_tile, int32_t, uint32_t, uint32_t*) { return not_available("si_mm_counts"); }
int si_batched_default_tuning(si_dataset*, si_tuning*) { return not_available("si_batched_default_tuning"); }
int si_batched_count(si_dataset*, const si_tuning*, uint32_t*) { return not_available("si_batched_count"); }
int si_batched_count_async(si_dataset*, const si_tuning*, si_buffer*) { return not_available("si_batched_count"); }
int si_batched_autotune(si_dataset*, double, si_tuning*) { return not_available("si_batched_autotune"); }
int si_edge_count(si_dataset*, const uint32_t*, const uint32_t*, uint64_t, const si_tuning*, uint32_t*, uint64_t*) {
    return not_available("si_edge_count");
}
int si_edge_count_async(si_dataset*, si_buffer*, si_buffer*, uint64_t, const si_tuning*, si_buffer*, si_buffer*) {
    return not_available("si_edge_count");
}

}  // extern "C"
