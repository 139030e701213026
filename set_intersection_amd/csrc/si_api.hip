// C-ABI entry points for counting: argument checks and technique dispatch.
#include <algorithm>

#include "si_internal.h"

using namespace si_impl;

namespace {

// Size ratio from which materialisation searches instead of merging.  The
// merge reads ratio x 4 B per element of the smaller set near the HBM rate;
// a search step past the LDS samples is a dependent line fetch, about a dozen
// of them per element for large sets.  Measured on MI355X (materialise, ms,
// merge / search): |small| = 2^20 at ratio 8: 0.047 / 0.057, 32: 0.076 /
// 0.099, 64: 0.125 / 0.125; 2^22 at 64: 0.415 / 0.432; 2^16 at 2048: 0.213 /
// 0.028.
constexpr uint64_t kSearchRatio = 64;

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

namespace si_impl {

int pair_count_dispatch(si_dataset* ds, uint32_t a, uint32_t b, int32_t technique, const si_params* params,
                        uint64_t* d) {
    switch (technique) {
        case SI_TECH_IP: return ip_pair_count(ds, a, b, params, d, nullptr);
        case SI_TECH_OBS: return obs_pair_count(ds, a, b, params, d, nullptr);
        case SI_TECH_HI: return hi_pair_count(ds, a, b, params, d, nullptr);
        case SI_TECH_BITMAP_DYNAMIC: return bitmap_pair_count(ds, a, b, params, d, nullptr);
        default: return not_available("si_pair_count");
    }
}

}  // namespace si_impl

extern "C" {

int si_pair_count_async(si_dataset* ds, uint32_t a, uint32_t b, int32_t technique, const si_params* params,
                        si_buffer* out, uint64_t offset) {
    SI_TRY(check_pair(ds, a, b));
    if (!out || offset + 8 > out->bytes || (offset & 7)) return fail(SI_ERR_ARG, "bad output buffer/offset");
    uint64_t* d = reinterpret_cast<uint64_t*>(static_cast<char*>(out->ptr) + offset);
    return pair_count_dispatch(ds, a, b, technique, params, d);
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
    // the staging slots and the output bound hold for strictly ascending sets
    // only (a run of duplicates would match more than min(|a|,|b|) times): a
    // dataset uploaded without validation is checked once, here
    if (!ds->validated) SI_TRY(validate_dataset(ds));
    // sizes at least kSearchRatio apart: the smaller set's elements are
    // searched in the larger one instead of merging both (unless a merge tile
    // shape is forced with SI_SISI_SHAPE)
    const uint64_t lo = std::min(ds->sizes[a], ds->sizes[b]), hi = std::max(ds->sizes[a], ds->sizes[b]);
    if (ds->ctx->sisi_shape == 0 && lo > 0 && hi / lo >= kSearchRatio)
        return obs_pair_materialize(ds, a, b, static_cast<uint32_t*>(out->ptr),
                                    static_cast<uint64_t*>(count_out->ptr));
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
    // A run of a few pairs (SISI datasets are k = 2) is spread over the whole
    // GPU pair by pair with the single-pair kernels.
    const bool self = A.id == B.id;
    const uint64_t la = A.end - A.start, lb = B.end - B.start;
    const uint64_t npairs = self ? la * (la ? la - 1 : 0) / 2 : la * lb;
    if (npairs <= 64 && technique != SI_TECH_BITMAP_NAIVE && technique != SI_TECH_MM) {
        for (uint32_t i = A.start; i < A.end; ++i) {
            for (uint32_t j = self ? i + 1 : B.start; j < B.end; ++j) {
                const uint64_t li = uint64_t(i) - uint64_t(A.id) * num_sets, lj = uint64_t(j) - uint64_t(B.id) * num_sets;
                const uint64_t n = num_sets;
                const uint64_t idx = self ? (n * (n - 1) / 2) - (n - li) * ((n - li) - 1) / 2 + lj - li - 1 : li * n + lj;
                uint32_t* slot = d + idx;
                switch (technique) {
                    case SI_TECH_IP: SI_TRY(ip_pair_count(ds, i, j, params, nullptr, slot)); break;
                    case SI_TECH_OBS: SI_TRY(obs_pair_count(ds, i, j, params, nullptr, slot)); break;
                    case SI_TECH_HI: SI_TRY(hi_pair_count(ds, i, j, params, nullptr, slot)); break;
                    case SI_TECH_BITMAP_DYNAMIC: SI_TRY(bitmap_pair_count(ds, i, j, params, nullptr, slot)); break;
                    default: return not_available("si_run_count");
                }
            }
        }
        return SI_OK;
    }
    switch (technique) {
        case SI_TECH_IP: return ip_run(ds, params, A, B, num_sets, d);
        case SI_TECH_OBS: return obs_run(ds, params, A, B, num_sets, d);
        case SI_TECH_HI: return hi_run(ds, params, A, B, num_sets, d);
        case SI_TECH_BITMAP_DYNAMIC: return bitmap_run(ds, params, A, B, num_sets, d);
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

int si_all_pairs_bitmap_async(si_dataset* ds, const si_params* params, si_buffer* counts) {
    if (!ds || !counts) return fail(SI_ERR_ARG, "null argument");
    const uint64_t slots = uint64_t(ds->k) * (ds->k > 0 ? ds->k - 1 : 0) / 2;
    if (counts->bytes < slots * 4) return fail(SI_ERR_CAPACITY, "counts buffer smaller than k(k-1)/2 entries");
    return bitmap_all_pairs(ds, params, static_cast<uint32_t*>(counts->ptr));
}

int si_all_pairs_bitmap(si_dataset* ds, const si_params* params, uint32_t* counts_host) {
    if (!ds || !counts_host) return fail(SI_ERR_ARG, "null argument");
    const uint64_t slots = uint64_t(ds->k) * (ds->k > 0 ? ds->k - 1 : 0) / 2;
    si_buffer* d = nullptr;
    SI_TRY(si_buffer_alloc(ds->ctx, std::max<uint64_t>(slots, 1) * 4, &d));
    int st = si_all_pairs_bitmap_async(ds, params, d);
    if (st == SI_OK) st = si_buffer_download(d, 0, counts_host, slots * 4);
    si_buffer_free(d);
    return st;
}

int si_mm_counts_async(si_dataset* ds, si_tile A, si_tile B, int32_t variant, uint32_t ld, si_buffer* counts) {
    if (!ds || !counts) return fail(SI_ERR_ARG, "null argument");
    SI_TRY(check_tile(ds, A));
    SI_TRY(check_tile(ds, B));
    if (variant < SI_MM_I8 || variant > SI_MM_FP4) return fail(SI_ERR_ARG, "unknown mm variant");
    const uint64_t ni = A.end - A.start, nj = B.end - B.start;
    if (ld < nj) return fail(SI_ERR_ARG, "ld smaller than B.len");
    const uint64_t need = ni ? ((ni - 1) * ld + nj) : 0;
    if (counts->bytes < need * 4) return fail(SI_ERR_CAPACITY, "counts buffer smaller than the tile matrix");
    return mm_tile_counts(ds, A, B, variant, ld, static_cast<uint32_t*>(counts->ptr));
}

int si_mm_counts(si_dataset* ds, si_tile A, si_tile B, int32_t variant, uint32_t ld, uint32_t* counts_host) {
    if (!ds || !counts_host) return fail(SI_ERR_ARG, "null argument");
    const uint64_t ni = A.end > A.start ? A.end - A.start : 0, nj = B.end > B.start ? B.end - B.start : 0;
    const uint64_t need = ni ? ((ni - 1) * uint64_t(ld) + nj) : 0;
    si_buffer* d = nullptr;
    SI_TRY(si_buffer_alloc(ds->ctx, std::max<uint64_t>(need, 1) * 4, &d));
    int st = si_mm_counts_async(ds, A, B, variant, ld, d);
    if (st == SI_OK) st = si_buffer_download(d, 0, counts_host, need * 4);
    si_buffer_free(d);
    return st;
}

int si_batched_default_tuning(si_dataset* ds, si_tuning* out) {
    if (!ds || !out) return fail(SI_ERR_ARG, "null argument");
    return batched_default_tuning(ds, out);
}

int si_batched_count_async(si_dataset* ds, const si_tuning* tuning, si_buffer* counts) {
    if (!ds || !counts) return fail(SI_ERR_ARG, "null argument");
    const uint64_t slots = uint64_t(ds->k) * (ds->k > 0 ? ds->k - 1 : 0) / 2;
    if (counts->bytes < slots * 4) return fail(SI_ERR_CAPACITY, "counts buffer smaller than k(k-1)/2 entries");
    return batched_count(ds, tuning, static_cast<uint32_t*>(counts->ptr));
}

int si_batched_count(si_dataset* ds, const si_tuning* tuning, uint32_t* counts_host) {
    if (!ds || !counts_host) return fail(SI_ERR_ARG, "null argument");
    const uint64_t slots = uint64_t(ds->k) * (ds->k > 0 ? ds->k - 1 : 0) / 2;
    si_buffer* d = nullptr;
    SI_TRY(si_buffer_alloc(ds->ctx, std::max<uint64_t>(slots, 1) * 4, &d));
    int st = si_batched_count_async(ds, tuning, d);
    if (st == SI_OK) st = si_buffer_download(d, 0, counts_host, slots * 4);
    si_buffer_free(d);
    return st;
}

int si_batched_autotune(si_dataset* ds, double budget_ms, si_tuning* out) {
    if (!ds || !out) return fail(SI_ERR_ARG, "null argument");
    return batched_autotune(ds, budget_ms, out);
}

int si_batched_autotune_seeded(si_dataset* ds, const si_tuning* seed, double budget_ms, si_tuning* out) {
    if (!ds || !seed || !out) return fail(SI_ERR_ARG, "null argument");
    return batched_autotune_seeded(ds, seed, budget_ms, out);
}

int si_edge_plan_create(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m, const si_tuning* tuning,
                        si_edge_plan** out) {
    if (!ds || !out || (m && (!u || !v))) return fail(SI_ERR_ARG, "null argument");
    *out = nullptr;
    return edge_plan_create(ds, u, v, m, tuning, out);
}

int si_edge_plan_free(si_edge_plan* plan) {
    if (plan) edge_plan_free(plan);
    return SI_OK;
}

int si_edge_count_async(si_edge_plan* plan, si_buffer* counts, si_buffer* total_out) {
    if (!plan || !counts || !total_out) return fail(SI_ERR_ARG, "null argument");
    if (total_out->bytes < 8) return fail(SI_ERR_CAPACITY, "total buffer smaller than 8 bytes");
    if (counts->bytes < plan_edges(plan) * 4) return fail(SI_ERR_CAPACITY, "counts buffer smaller than m entries");
    return edge_count(plan, static_cast<uint32_t*>(counts->ptr), static_cast<uint64_t*>(total_out->ptr));
}

int si_edge_autotune(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m, double budget_ms,
                     si_tuning* out) {
    if (!ds || !out || (m && (!u || !v))) return fail(SI_ERR_ARG, "null argument");
    return edge_autotune(ds, u, v, m, budget_ms, out);
}

int si_edge_count_plain_async(si_edge_plan* plan, si_buffer* counts, si_buffer* total_out) {
    if (!plan || !counts || !total_out) return fail(SI_ERR_ARG, "null argument");
    if (total_out->bytes < 8) return fail(SI_ERR_CAPACITY, "total buffer smaller than 8 bytes");
    if (counts->bytes < plan_edges(plan) * 4) return fail(SI_ERR_CAPACITY, "counts buffer smaller than m entries");
    return edge_count_plain(plan, static_cast<uint32_t*>(counts->ptr), static_cast<uint64_t*>(total_out->ptr));
}

int si_edge_count(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m, const si_tuning* tuning,
                  uint32_t* counts_host, uint64_t* total) {
    if (!ds || (m && (!u || !v || !counts_host)) || !total) return fail(SI_ERR_ARG, "null argument");
    si_edge_plan* plan = nullptr;
    SI_TRY(si_edge_plan_create(ds, u, v, m, tuning, &plan));
    si_buffer* dc = nullptr;
    si_buffer* dt = nullptr;
    int st = si_buffer_alloc(ds->ctx, std::max<uint64_t>(m, 1) * 4, &dc);
    if (st == SI_OK) st = si_buffer_alloc(ds->ctx, 8, &dt);
    if (st == SI_OK) st = si_edge_count_async(plan, dc, dt);
    if (st == SI_OK) st = si_buffer_download(dc, 0, counts_host, m * 4);
    if (st == SI_OK) st = si_buffer_download(dt, 0, total, 8);
    si_buffer_free(dc);
    si_buffer_free(dt);
    si_edge_plan_free(plan);
    return st;
}

}  // extern "C"
