// Hash-intersect ("hi") kernels for gfx950.
//
// Reference behaviour (src/hi.cu, adapted there from H-INDEX triangle
// counting): a block (or warp) hashes set a into buckets `x % buckets` held in
// global memory with per-bucket counts in shared memory (hi.cu:379-391,
// 319-331), then every element of b probes its bucket with a linear scan
// (linearSearch, hi.cu:11-32; probes hi.cu:403-407).  --path hashes and probes
// merge-path partitions (hi.cu:420-479).
//
// MI355X design:
//  * Open-addressing tables (multiplicative hash, linear probing, load <= 1/2)
//    sized per set: in LDS when they fit, otherwise in a per-group region of
//    global memory.  Capacity is derived from |a|, so no bucket can overflow
//    (the reference never checks bucketSize, hi.cu:329-330).  0xFFFFFFFF is
//    the empty-slot marker; the value itself is tracked by a flag.
//  * SISI cuts the pair into merge-path segments (A-first tie order, so a
//    matching pair never straddles two segments once the B range includes the
//    B element that follows it) and each workgroup hashes the SMALLER side of
//    its segment in LDS and probes the other side.
//  * MISI runs keep the reference's ownership (workgroup, or one wave with
//    --warp, per row a) and reduce each pair with wave shuffles.
#include "si_device.h"
#include "si_internal.h"

namespace {

using namespace si_dev;

constexpr uint32_t EMPTY = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t hslot(uint32_t x, uint32_t logTS) {
    return (x * 0x9E3779B1u) >> (32 - logTS);
}

__device__ __forceinline__ void table_insert(uint32_t* tab, uint32_t mask, uint32_t logTS, uint32_t x, uint32_t* hasMax) {
    if (x == EMPTY) {
        *hasMax = 1;
        return;
    }
    uint32_t s = hslot(x, logTS);
    while (true) {
        const uint32_t prev = atomicCAS(&tab[s], EMPTY, x);
        if (prev == EMPTY || prev == x) return;
        s = (s + 1) & mask;
    }
}

__device__ __forceinline__ uint32_t table_find(const uint32_t* tab, uint32_t mask, uint32_t logTS, uint32_t x,
                                               uint32_t hasMax) {
    if (x == EMPTY) return hasMax;
    uint32_t s = hslot(x, logTS);
    while (true) {
        const uint32_t k = tab[s];
        if (k == x) return 1;
        if (k == EMPTY) return 0;
        s = (s + 1) & mask;
    }
}

__device__ __forceinline__ uint32_t log2_ceil(uint64_t v) {
    uint32_t l = 0;
    while ((uint64_t(1) << l) < v) ++l;
    return l;
}

// ---------------------------------------------------------------- SISI ----
template <int NT, int SEG>
__global__ __launch_bounds__(NT) void k_hash_pair(const uint32_t* __restrict__ base, uint64_t oA, uint64_t nA,
                                                  uint64_t oB, uint64_t nB, uint64_t* __restrict__ partials,
                                                  uint32_t* __restrict__ done, uint64_t* __restrict__ out64,
                                                  uint32_t* __restrict__ out32) {
    // the hashed side is the smaller one: <= (SEG + 1) / 2 elements, so SEG
    // slots keep the load factor <= 1/2
    static_assert((SEG & (SEG - 1)) == 0, "SEG must be a power of two");
    constexpr uint32_t LOG = 31 - __builtin_clz(SEG);
    __shared__ uint32_t tab[1u << LOG];
    __shared__ uint64_t seg[2];
    __shared__ uint32_t hasMax;
    __shared__ uint64_t red[NT / WAVE];
    const uint64_t n = nA + nB;
    const uint64_t dBeg = uint64_t(blockIdx.x) * SEG;
    const uint64_t dEnd = dBeg + SEG < n ? dBeg + SEG : n;
    if (threadIdx.x < WAVE) {
        const uint64_t c = corank_halfwave(base + oA, nA, base + oB, nB, lane_id() < 32 ? dBeg : dEnd);
        if (lane_id() == 0) seg[0] = c;
        if (lane_id() == 32) seg[1] = c;
    }
    for (uint32_t i = threadIdx.x; i < (1u << LOG); i += NT) tab[i] = EMPTY;
    if (threadIdx.x == 0) hasMax = 0;
    __syncthreads();
    const uint64_t ia = seg[0], iaE = seg[1];
    const uint64_t ib = dBeg - ia;
    uint64_t ibE = dEnd - iaE;
    if (ibE < nB) ++ibE;  // include the B element that follows the segment
    const uint64_t na = iaE - ia, nb = ibE - ib;
    // hash the smaller side
    const bool hashA = na <= nb;
    const uint32_t* H = base + (hashA ? oA + ia : oB + ib);
    const uint32_t* Q = base + (hashA ? oB + ib : oA + ia);
    const uint32_t nh = uint32_t(hashA ? na : nb), nq = uint32_t(hashA ? nb : na);
    const uint32_t mask = (1u << LOG) - 1;
    for (uint32_t i = threadIdx.x; i < nh; i += NT) table_insert(tab, mask, LOG, H[i], &hasMax);
    __syncthreads();
    uint64_t c = 0;
    const uint32_t hm = hasMax;
    for (uint32_t i = threadIdx.x; i < nq; i += NT) c += table_find(tab, mask, LOG, Q[i], hm);
    grid_fold_u64<NT>(c, partials, done, out64, out32, red);
}

// ---------------------------------------------------------------- MISI ----
// LDS_SLOTS: table capacity in LDS; larger rows use gtab + blockIdx * gslots.
template <int NT, int LDS_SLOTS, bool SELF>
__global__ __launch_bounds__(NT) void k_hash_run(const uint32_t* __restrict__ base, const uint64_t* __restrict__ offsets,
                                                 RunArgs r, uint32_t* __restrict__ counts, uint32_t* __restrict__ gtab,
                                                 uint32_t gslots) {
    __shared__ uint32_t ltab[LDS_SLOTS];
    __shared__ uint32_t hasMax;
    for (uint32_t a = r.aStart + blockIdx.x; a < r.aEnd; a += gridDim.x) {
        const uint64_t oA = offsets[a], nA = offsets[a + 1] - oA;
        if (nA == 0) continue;
        const uint32_t logTS = log2_ceil(nA * 2 < 64 ? 64 : nA * 2);
        const uint32_t TS = 1u << logTS, mask = TS - 1;
        uint32_t* tab = (TS <= uint32_t(LDS_SLOTS)) ? ltab : gtab + uint64_t(blockIdx.x) * gslots;
        for (uint32_t i = threadIdx.x; i < TS; i += NT) tab[i] = EMPTY;
        if (threadIdx.x == 0) hasMax = 0;
        __threadfence_block();
        __syncthreads();
        for (uint64_t i = threadIdx.x; i < nA; i += NT) table_insert(tab, mask, logTS, base[oA + i], &hasMax);
        __threadfence_block();
        __syncthreads();
        const uint32_t hm = hasMax;
        for (uint32_t b = SELF ? a + 1 : r.bStart; b < r.bEnd; ++b) {
            const uint64_t oB = offsets[b], nB = offsets[b + 1] - oB;
            uint32_t c = 0;
            for (uint64_t i = threadIdx.x; i < nB; i += NT) c += table_find(tab, mask, logTS, base[oB + i], hm);
            c = wave_sum_u32(c);
            if (lane_id() == 0 && c) atomicAdd(&counts[run_index<SELF>(r, a, b)], c);
        }
        __syncthreads();
    }
}

constexpr int PAIR_NT = 256;
constexpr int PAIR_SEG = 8192;

}  // namespace

namespace si_impl {

int hi_pair_count(si_dataset* ds, uint32_t a, uint32_t b, const si_params*, uint64_t* d_out, uint32_t* d_out32) {
    si_context* ctx = ds->ctx;
    const uint64_t oA = ds->offsets[a], nA = ds->sizes[a], oB = ds->offsets[b], nB = ds->sizes[b];
    if (nA == 0 || nB == 0) {
        return zero_outputs(ctx, d_out, d_out32);
    }
    const uint64_t nseg = (nA + nB + PAIR_SEG - 1) / PAIR_SEG;
    if (nseg > 0x7FFFFFFFull) return fail(SI_ERR_RANGE, "pair too large for the hash kernel grid");
    void* ws = nullptr;
    SI_TRY(small_workspace(ctx, 256 + nseg * 8, &ws));
    KernelScope ks(ctx, "hi_sisi_count");
    hipLaunchKernelGGL((k_hash_pair<PAIR_NT, PAIR_SEG>), dim3(uint32_t(nseg)), dim3(PAIR_NT), 0, ctx->stream,
                       ds->d_elems, oA, nA, oB, nB, reinterpret_cast<uint64_t*>(static_cast<char*>(ws) + 256),
                       static_cast<uint32_t*>(ws), d_out, d_out32);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int hi_run(si_dataset* ds, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts) {
    si_context* ctx = ds->ctx;
    const bool self = (A.id == B.id);
    RunArgs r{A.start, A.end, A.id, B.start, B.end, B.id, n};
    const bool warp = p && p->warp;
    uint32_t maxA = 0;
    for (uint32_t i = A.start; i < A.end; ++i) maxA = std::max(maxA, ds->sizes[i]);
    uint64_t ts = 64;
    while (ts < 2ull * maxA) ts <<= 1;
    const uint32_t ldsSlots = warp ? 4096 : 16384;
    uint64_t grid = (p && p->blocks) ? p->blocks : uint64_t(ctx->num_cus) * (warp ? 32 : 8);
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, A.end - A.start));
    uint32_t gslots = 0;
    uint32_t* gtab = nullptr;
    if (ts > ldsSlots) {
        // global tables: cap the workspace at 2 GiB by reducing the grid
        const uint64_t budget = 2ull << 30;
        grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, budget / (ts * 4)));
        SI_TRY(ctx->ws_aux.ensure(grid * ts * 4));
        gtab = ctx->ws_aux.as<uint32_t>();
        gslots = uint32_t(ts);
    }
    KernelScope ks(ctx, warp ? "hi_run_warp" : "hi_run_block");
#define LAUNCH(NT, SL, S)                                                                                        \
    hipLaunchKernelGGL((k_hash_run<NT, SL, S>), dim3(uint32_t(grid)), dim3(NT), 0, ctx->stream, ds->d_elems, \
                       ds->d_offsets, r, d_counts, gtab, gslots)
    if (warp) {
        if (self) LAUNCH(64, 4096, true); else LAUNCH(64, 4096, false);
    } else {
        if (self) LAUNCH(256, 16384, true); else LAUNCH(256, 16384, false);
    }
#undef LAUNCH
    SI_HIP(hipGetLastError());
    return SI_OK;
}

}  // namespace si_impl
