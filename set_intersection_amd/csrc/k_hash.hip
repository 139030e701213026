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
//  * SISI is tiled by value: the universe is cut into buckets of 2^bits
//    values (about 4096 elements of the larger set each) and a workgroup
//    takes a run of consecutive buckets.  Per bucket it inserts the bucket's
//    elements of the SMALLER set into an LDS table sized for them (16-byte
//    groups of a fill counter and three key slots, at least eight slots per
//    element where the table allows) and probes it with the bucket's
//    elements of the larger set; both sets stream once from HBM in order.
//    A key is the value's low `bits` bits (0xFFFFFFFF marks an empty slot).
//    An insert claims a slot with one LDS atomic on its home group's counter
//    (keys are unique, so no compare-and-swap) and passes on to the next
//    group when the group is full; a probe reads one group and compares
//    three keys, and only a full group sends it further.  With --path the
//    pair is cut into merge-path segments instead (k_path.hip).
//  * MISI runs keep the reference's ownership (workgroup, or one wave with
//    --warp, per row a) and reduce each pair with wave shuffles.
#include <cmath>

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
// One workgroup per run of `per` buckets [v0, v1) of 2^bits values; dynamic
// LDS holds the bucket bounds of both sets.  S (the smaller set) is hashed,
// Q probes.  A thread holds up to U elements of each side of the current
// bucket in registers, and the next bucket's are loaded while the current one
// is processed.  A bucket with more than TSMAX/2 elements of S is hashed in
// pieces of TSMAX/2, each probed by all of the bucket's Q elements; the
// bucket width keeps the expected S elements per bucket at 3072 or fewer.
template <int NT, int U, int TSMAX>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_hash_buckets(const uint32_t* __restrict__ S, uint64_t nS,
                                                     const uint32_t* __restrict__ Q, uint64_t nQ, uint32_t bits,
                                                     uint64_t nb, uint32_t per, uint64_t* __restrict__ partials,
                                                     uint32_t* __restrict__ done, uint64_t* __restrict__ out64,
                                                     uint32_t* __restrict__ out32, uint32_t prio) {
    static_assert(TSMAX / 2 <= NT * U, "the first piece of a bucket's S elements comes from registers");
    extern __shared__ uint64_t bnd[];  // S bounds [0, nbd), Q bounds [nbd, 2 nbd)
    __shared__ uint4 tab4[TSMAX / 4];
    __shared__ uint64_t red[NT / WAVE];
    uint32_t* tab = reinterpret_cast<uint32_t*>(tab4);
    const uint32_t low = (1u << bits) - 1u;
    const uint64_t v0 = uint64_t(blockIdx.x) * per, v1 = v0 + per < nb ? v0 + per : nb;
    uint64_t c = 0;
    if (v0 < v1) {
        const uint32_t nbd = uint32_t(v1 - v0) + 1;
        for (uint32_t t = threadIdx.x; t < 2 * nbd; t += NT) {
            const bool isQ = t >= nbd;
            const uint64_t key = (v0 + (isQ ? t - nbd : t)) << bits;  // first value of the bucket
            const uint32_t* P = isQ ? Q : S;
            uint64_t lo = 0, hi = isQ ? nQ : nS;  // first position with a value >= key
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (uint64_t(P[mid]) < key) lo = mid + 1; else hi = mid;
            }
            bnd[t] = lo;
        }
        __syncthreads();
        auto work = [&](uint32_t k) {  // the next bucket with elements on both sides
            while (k + 1 < nbd && (bnd[k] == bnd[k + 1] || bnd[nbd + k] == bnd[nbd + k + 1])) ++k;
            return k;
        };
        // a bucket's elements of each side, NT x U at a time: a scalar base
        // (the bucket's first element) and a 32-bit byte offset per load,
        // clamped to the side's last element
        auto fetch = [&](uint32_t k, uint32_t (&xs)[U], uint32_t (&xq)[U]) {
            const uint64_t s0 = bnd[k], s1 = bnd[k + 1], q0 = bnd[nbd + k], q1 = bnd[nbd + k + 1];
            const char* sb = reinterpret_cast<const char*>(S + uniform_u64(s0));
            const char* qb = reinterpret_cast<const char*>(Q + uniform_u64(q0));
            // a bucket holds at most 2^30 elements (bits <= 30): 4 (n - 1) fits 32 bits
            const uint32_t sl = 4u * uniform_u32(uint32_t(s1 - s0) - 1u), ql = 4u * uniform_u32(uint32_t(q1 - q0) - 1u);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t o = 4u * (threadIdx.x + uint32_t(u) * NT);
                xs[u] = *reinterpret_cast<const uint32_t*>(sb + (o < sl ? o : sl));
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t o = 4u * (threadIdx.x + uint32_t(u) * NT);
                xq[u] = *reinterpret_cast<const uint32_t*>(qb + (o < ql ? o : ql));
            }
        };
        constexpr uint64_t CAP = uint64_t(NT) * U;
        constexpr uint32_t PIECE = TSMAX / 2;
        uint32_t hits = 0;  // this thread's matches from its registers (at most U per bucket)
        uint32_t xs[U], xq[U], ys[U], yq[U];
        uint32_t k = work(0);
        if (k + 1 < nbd) fetch(k, xs, xq);
        uint32_t step = 0;
        while (k + 1 < nbd) {
            if (prio) set_wave_prio((step++ + blockIdx.x) & 3u);
            const uint32_t kn = work(k + 1);
            if (kn + 1 < nbd) fetch(kn, ys, yq);
            const uint64_t s0 = bnd[k], s1 = bnd[k + 1], q0 = bnd[nbd + k], q1 = bnd[nbd + k + 1];
            const uint32_t qn = uniform_u32(uint32_t(q1 - q0 < CAP ? q1 - q0 : CAP));
            for (uint64_t p0 = s0; p0 < s1; p0 += PIECE) {
                const uint32_t m = uint32_t(s1 - p0 < PIECE ? s1 - p0 : PIECE);
                // 2^lg groups of 16 bytes: a fill counter and three key slots.
                // At least 8m slots where TSMAX allows (most probes end in
                // their first group), else 1.5 m (m <= TSMAX / 2)
                constexpr uint32_t LGMAX = uint32_t(__builtin_ctz(uint32_t(TSMAX)));
                uint32_t lg = 32u - uint32_t(__builtin_clz(8 * m - 1));
                lg = lg < LGMAX ? lg : LGMAX;
                lg = (lg < 6 ? 6 : lg) - 2;
                const uint32_t gmask = (1u << lg) - 1u;
                for (uint32_t i = threadIdx.x; i <= gmask; i += NT) tab4[i] = make_uint4(0u, EMPTY, EMPTY, EMPTY);
                __syncthreads();
                // key: the value's low bits (below EMPTY); home group from a
                // 24-bit multiplicative hash (full-rate multiply)
                auto home = [&](uint32_t key) { return uint32_t(__umul24(key, 0x9E3779u)) >> (32 - lg); };  // HIP's __umul24 returns int
                // claim a slot with one atomic on the group's counter; a full
                // group passes the key on to the next (keys are unique)
                auto insert = [&](uint32_t key, uint32_t g) {
                    while (true) {
                        const uint32_t c = atomicAdd(&tab[4u * g], 1u);
                        if (c < 3u) {
                            tab[4u * g + 1u + c] = key;
                            return;
                        }
                        g = (g + 1u) & gmask;
                    }
                };
                if (p0 == s0) {
                    // the first piece from registers: every element's first
                    // claim, then the stores (a claim past slot 3 walks on)
#pragma unroll
                    for (int b = 0; b < U; b += 4) {
                        if (uint32_t(b) * NT >= m) break;
                        uint32_t g[4], c[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const bool v = threadIdx.x + uint32_t(b + j) * NT < m;
                            g[j] = home(xs[b + j] & low);
                            c[j] = v ? atomicAdd(&tab[4u * g[j]], 1u) : 3u;
                        }
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const bool v = threadIdx.x + uint32_t(b + j) * NT < m;
                            if (v && c[j] < 3u) tab[4u * g[j] + 1u + c[j]] = xs[b + j] & low;
                            else if (v) insert(xs[b + j] & low, (g[j] + 1u) & gmask);
                        }
                    }
                } else {
                    for (uint64_t i = p0 + threadIdx.x; i < p0 + m; i += NT) {
                        const uint32_t key = S[i] & low;
                        insert(key, home(key));
                    }
                }
                __syncthreads();
                // a key is present if its group, or a full group after it
                // up to the first group that is not full, holds it
                auto probe_rest = [&](uint32_t key, uint32_t g) {
                    while (true) {
                        g = (g + 1u) & gmask;
                        const uint4 t = tab4[g];
                        if (t.y == key || t.z == key || t.w == key) return true;
                        if (t.x < 3u) return false;
                    }
                };
                // four probes at a time: their group reads are in flight together
#pragma unroll
                for (int b = 0; b < U; b += 4) {
                    if (uint32_t(b) * NT >= qn) break;
                    uint32_t key[4], g[4];
                    uint4 t[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        key[j] = xq[b + j] & low;
                        g[j] = home(key[j]);
                        t[j] = tab4[g[j]];
                    }
                    bool more[4], anyMore = false;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const bool v = threadIdx.x + uint32_t(b + j) * NT < qn;
                        const bool hit = v && (t[j].y == key[j] || t[j].z == key[j] || t[j].w == key[j]);
                        hits += hit ? 1u : 0u;
                        more[j] = v && !hit && t[j].x >= 3u;
                        anyMore = anyMore || more[j];
                    }
                    if (__builtin_amdgcn_ballot_w64(anyMore)) {  // rare: a full home group
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (more[j] && probe_rest(key[j], g[j])) ++hits;
                    }
                }
                for (uint64_t i = q0 + CAP + threadIdx.x; i < q1; i += NT) {
                    const uint32_t key = Q[i] & low, g = home(key);
                    const uint4 t = tab4[g];
                    c += (t.y == key || t.z == key || t.w == key || (t.x >= 3u && probe_rest(key, g))) ? 1u : 0u;
                }
                __syncthreads();  // probes done before the table is cleared again
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                xs[u] = ys[u];
                xq[u] = yq[u];
            }
            k = kn;
        }
        c += hits;
    }
    if (prio) set_wave_prio(0);
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
constexpr int PAIR_TSMAX = 8192;  // 32 KiB table: four workgroups per CU

}  // namespace

namespace si_impl {

int hi_pair_count(si_dataset* ds, uint32_t a, uint32_t b, const si_params* p, uint64_t* d_out, uint32_t* d_out32) {
    if (p && p->path) return path_pair_count(ds, a, b, SI_TECH_HI, p, d_out, d_out32);
    si_context* ctx = ds->ctx;
    uint32_t s = a, q = b;  // hash the smaller set, probe with the larger
    if (ds->sizes[s] > ds->sizes[q]) std::swap(s, q);
    const uint64_t nS = ds->sizes[s], nQ = ds->sizes[q];
    if (nS == 0 || nQ == 0) return zero_outputs(ctx, d_out, d_out32);
    // buckets of 2^bits values holding about 4096 elements of the larger set
    // (what the workgroup's registers take at once)
    const uint64_t vmax = uint64_t(std::max(ds->max_elem, ds->universe));
    uint32_t lu = 0, ln = 0;
    while ((uint64_t(1) << lu) <= vmax) ++lu;  // values fit in lu bits
    while ((uint64_t(1) << (ln + 1)) <= nQ) ++ln;
    int bits = int(lu) - int(ln) + 12;
    // expected elements of the hashed set per bucket at most 3072, so that a
    // bucket almost never needs a second piece of 4096 (table load <= 1/2)
    while (bits > 6 && double(nS) * std::ldexp(1.0, bits) > 3072.0 * (double(vmax) + 1.0)) --bits;
    bits = bits < 6 ? 6 : (bits > 30 ? 30 : bits);
    const uint64_t nb = (vmax >> bits) + 1;
    const uint64_t maxGrid = (p && p->blocks) ? p->blocks : uint64_t(ctx->num_cus) * 4;
    uint64_t per = (nb + maxGrid - 1) / maxGrid;
    const uint64_t perMax = 255;  // bucket bounds: 4 KiB of dynamic LDS beside the 32 KiB table
    if (per > perMax) per = perMax;
    const uint64_t g = (nb + per - 1) / per;
    if (g > 0x7FFFFFFFull) return fail(SI_ERR_RANGE, "pair too large for the hash kernel grid");
    const size_t lds = size_t(2) * (per + 1) * 8;
    void* ws = nullptr;
    SI_TRY(small_workspace(ctx, 256 + g * 8, &ws));
    KernelScope ks(ctx, "hi_sisi_count");
    hipLaunchKernelGGL((k_hash_buckets<PAIR_NT, 16, PAIR_TSMAX>), dim3(uint32_t(g)), dim3(PAIR_NT), lds, ctx->stream,
                       ds->d_elems + ds->offsets[s], nS, ds->d_elems + ds->offsets[q], nQ, uint32_t(bits), nb,
                       uint32_t(per), reinterpret_cast<uint64_t*>(static_cast<char*>(ws) + 256),
                       static_cast<uint32_t*>(ws), d_out, d_out32, uint32_t(ctx->sisi_prio));
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int hi_run(si_dataset* ds, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts) {
    if (p && p->path) return path_run(ds, SI_TECH_HI, p, A, B, n, d_counts);
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
