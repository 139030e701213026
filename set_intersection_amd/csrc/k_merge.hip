// Merge-path ("intersect path", technique `ip`) kernels for gfx950.
//
// Reference behaviour: src/ip.cu:141-177 (driver), findDiagonals
// (include/helpers.cuh:74-166) and intersectPath (ip.cu:207-245): the merge
// path of a pair is cut into equal-length partitions, one per block, and each
// thread merges `vt` elements of its block's partition, counting equal pairs.
//
// MI355X design (not a translation):
//  * One workgroup owns one contiguous SEGMENT of the merge path (SISI) or one
//    whole pair (MISI).  Only the segment ends need a global co-rank search;
//    it is done by half-waves with a 32-ary search (ceil(log32 n) dependent
//    rounds instead of log2 n).
//  * SISI (k_sisi_stream): the segment is streamed in tiles of up to
//    T = NT*VT merged elements (VT odd, so that lanes reading about VT/2
//    apart spread over the LDS banks) through two circular LDS windows of
//    W ~ T/2 elements, which doubles the resident waves per CU over
//    T-element windows.  A thread's split is kept only if it consumes at most
//    W-1 elements of a side; feasible splits form a prefix of the threads and
//    the tile is cut after them (for random data almost never).  Refill loads
//    (aligned 128-bit, one read of every input element) are issued before the
//    tile's merge and committed to LDS after it.
//  * Per tile every thread finds its split by a branch-free binary search in
//    LDS at the END of its diagonal range; its start is its neighbour's end.
//    The merge loop re-reads both heads each step (two independent LDS reads,
//    five VALU), and in count mode matches are counted per wave with
//    ballot + scalar popcount.
//  * Ties put the A element first, so an A element is matched against the B
//    element that is next in line when it is merged; no partition can split a
//    matching pair and no "count augmentation" (helpers.cuh:121-123) is needed.
//  * Counts are reduced with wave64 shuffles, one value per workgroup; the
//    last workgroup to finish folds the per-segment partials (no atomics on
//    the result, no memset before a launch).
//  * Materialisation writes each segment's matches, in order, to a staging
//    slot with buffer stores that the hardware discards for non-matches, then
//    places the slot at its prefix offset in the same launch (decoupled
//    look-back over per-segment counts).
#include <algorithm>

#include "si_device.h"
#include "si_internal.h"
#include "si_merge.h"

namespace {

using namespace si_dev;

// ---------------------------------------------------------------- SISI ----
// One workgroup per merge-path segment of the pair (A, B).
// Count mode: partials[s] holds segment s's count; the last workgroup to
// finish (ticket done[0]) folds them into out / out32 and resets done[0].
// Materialise mode: segments are numbered by a second ticket (done[1]), so
// the predecessors of every segment are already running.  A segment stages
// its matches in its slot of `stage`, publishes its count to state[s], adds
// up its predecessors' counts (decoupled look-back, 64 predecessors per round:
// an inclusive prefix ends the walk) and copies the slot to list[prefix ...).  The last segment writes
// the total to out; the last workgroup to finish clears state and resets
// done[0] and done[1].
// Start of segment s when the nseg segment lengths rise linearly from
// (1 - e) to (1 + e) times the mean, e = ramp / 256: n s ((1 - e)(nseg - 1) +
// e (s - 1)) / (nseg (nseg - 1)), split so no product passes 64 bits:
// q and den are below 2^8 nseg^2, so (n % den) q < 2^16 nseg^4 needs
// nseg < 2^12 (ramp < 256).  Monotone, 0 at s = 0 and n at s = nseg.
__host__ __device__ inline uint64_t ramp_start(uint64_t n, uint32_t nseg, uint32_t ramp, uint32_t s) {
    if (nseg < 2 || s == 0) return s == 0 ? 0 : n;
    if (s >= nseg) return n;
    const uint64_t q = uint64_t(s) * (uint64_t(256 - ramp) * (nseg - 1) + uint64_t(ramp) * (s - 1));
    const uint64_t den = uint64_t(256) * nseg * (nseg - 1);
    return (n / den) * q + ((n % den) * q) / den;
}

template <int NT, int VT, int WN, bool MAT, bool SEL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_sisi_stream(const uint32_t* __restrict__ base, uint64_t oA, uint64_t nA,
                                                    uint64_t oB, uint64_t nB, uint64_t seg_len, uint32_t nseg,
                                                    uint64_t* __restrict__ partials, uint32_t* __restrict__ done,
                                                    uint64_t* __restrict__ out, uint32_t* __restrict__ out32,
                                                    uint32_t* __restrict__ stage, uint64_t stage_cap,
                                                    uint32_t* __restrict__ list, uint64_t* __restrict__ state,
                                                    uint32_t prio, uint32_t ramp) {
    __shared__ StreamSmem<NT, VT, WN, MAT ? 1024 : 1> sm;
    __shared__ uint32_t shTicket;
    __shared__ uint64_t shPrefix;
    uint32_t s = blockIdx.x;
    if (MAT) {
        if (threadIdx.x == 0) shTicket = atomicAdd(&done[1], 1u);
        __syncthreads();
        s = shTicket;
    }
    const uint64_t n = nA + nB;
    uint64_t dBeg = uint64_t(s) * seg_len;
    uint64_t dEnd = (dBeg + seg_len < n) ? dBeg + seg_len : n;
    if (ramp) {  // segment lengths rising linearly with the ticket
        dBeg = ramp_start(n, nseg, ramp, s);
        dEnd = ramp_start(n, nseg, ramp, s + 1);
    }

    if (threadIdx.x < WAVE) {
        const uint64_t dd = (lane_id() < 32) ? dBeg : dEnd;
        const uint64_t c = corank_halfwave(base + oA, nA, base + oB, nB, dd);
        if (lane_id() == 0) sm.seg[0] = c;
        if (lane_id() == 32) sm.seg[1] = c;
    }
    __syncthreads();
    const uint64_t ia = sm.seg[0], iaE = sm.seg[1];
    const uint64_t ib = dBeg - ia, ibE = dEnd - iaE;

    uint32_t* slot = MAT ? stage + uint64_t(s) * stage_cap : nullptr;
    const uint64_t c =
        merge_segment_stream<NT, VT, WN, MAT, SEL, MAT ? 1024 : 1>(base, oA, nA, oB, nB, ia, ib, iaE, ibE, sm, slot,
                                                             prio ? s : kNoPrio);
    const uint64_t total = block_sum_u64<NT>(c, sm.red);

    if (MAT) {
        if (threadIdx.x < WAVE) {
            const uint64_t pre = lookback_exclusive(state, s, total);
            if (lane_id() == 0) {
                shPrefix = pre;
                if (s == nseg - 1 && out) *out = pre + total;
            }
        }
        __syncthreads();
        // place the slot: eight loads in flight per thread
        const uint64_t pre = shPrefix;
        constexpr int CU = 8;
        for (uint64_t i0 = threadIdx.x; i0 < total; i0 += uint64_t(NT) * CU) {
            uint32_t v[CU];
#pragma unroll
            for (int u = 0; u < CU; ++u) {
                const uint64_t i = i0 + uint64_t(u) * NT;
                v[u] = i < total ? slot[i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < CU; ++u) {
                const uint64_t i = i0 + uint64_t(u) * NT;
                if (i < total) list[pre + i] = v[u];
            }
        }
        if (threadIdx.x == 0) shTicket = atomicAdd(&done[0], 1u);
        __syncthreads();
        if (shTicket == nseg - 1) {  // every segment has finished its look-back
            for (uint32_t i = threadIdx.x; i < nseg; i += NT)
                __hip_atomic_store(&state[i], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (threadIdx.x == 0) {
                __hip_atomic_store(&done[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&done[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        return;
    }

    // count mode: the last workgroup folds the partials
    if (threadIdx.x == 0) partials[s] = total;
    __shared__ uint32_t isLast;
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t ticket = atomicAdd(done, 1u);
        isLast = (ticket == nseg - 1);
    }
    __syncthreads();
    if (!isLast) return;
    __threadfence();
    uint64_t acc = 0;
    for (uint32_t i = threadIdx.x; i < nseg; i += NT) acc += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    acc = block_sum_u64<NT>(acc, sm.red);
    if (threadIdx.x == 0) {
        if (out) *out = acc;
        if (out32) *out32 = uint32_t(acc);
        *done = 0;
    }
}

// ---------------------------------------------------------------- MISI ----
// One 64-thread workgroup per pair, grid-stride over the pairs of a run.
template <int VT, bool SELF>
__global__ __launch_bounds__(WAVE) void k_run_merge(const uint32_t* __restrict__ base,
                                                    const uint64_t* __restrict__ offsets, uint32_t aStart,
                                                    uint32_t aLen, uint32_t aId, uint32_t bStart, uint32_t bLen,
                                                    uint32_t bId, uint32_t n, uint64_t npairs,
                                                    uint32_t* __restrict__ counts) {
    __shared__ MergeSmem<WAVE, VT> sm;
    for (uint64_t p = blockIdx.x; p < npairs; p += gridDim.x) {
        uint32_t li, lj;
        if (SELF) {
            tri_decode(p, aLen, li, lj);
        } else {
            li = uint32_t(p / bLen);
            lj = uint32_t(p % bLen);
        }
        const uint32_t a = aStart + li, b = bStart + lj;
        const uint64_t oA = offsets[a], nA = offsets[a + 1] - oA;
        const uint64_t oB = offsets[b], nB = offsets[b + 1] - oB;
        uint64_t c = 0;
        if (nA && nB) c = merge_segment<WAVE, VT, false>(base, oA, nA, oB, nB, 0, 0, nA, nB, sm, nullptr);
        c = wave_sum_u64(c);
        if (threadIdx.x == 0) {
            const uint64_t ia = uint64_t(a) - uint64_t(aId) * n, ib = uint64_t(b) - uint64_t(bId) * n;
            const uint64_t idx = SELF ? tri_index(n, ia, ib) : ia * n + ib;
            counts[idx] = uint32_t(c);
        }
        __syncthreads();
    }
}

constexpr int RUN_VT = 8;

}  // namespace

namespace si_impl {

// Tile shapes of the SISI merge kernel (threads x elements per thread, window
// W) and workgroups per CU; 0 and 6 are chosen automatically (auto_shape), the
// others are selectable through si_params.reserved[0] (count) or
// SI_SISI_SHAPE (materialisation) for tuning sweeps; shape 1 reads one head
// per merge step.
struct SisiShape {
    int nt, vt, per_cu;
};
static const SisiShape kShapes[] = {{256, 31, 4}, {256, 31, 4}, {128, 31, 8}, {256, 15, 8}, {128, 15, 16},
                                    {512, 15, 4}, {256, 15, 4}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

static void sisi_geometry(si_context* ctx, uint64_t n, const si_params* p, const SisiShape& sh, uint32_t& nseg,
                          uint64_t& seg_len) {
    const uint64_t T = uint64_t(sh.nt) * sh.vt;
    uint64_t maxSeg = uint64_t(ctx->num_cus) * sh.per_cu;
    if (p && p->blocks) maxSeg = p->blocks;
    uint64_t want = (n + T - 1) / T;
    if (want < 1) want = 1;
    if (want > maxSeg) want = maxSeg;
    seg_len = (n + want - 1) / want;
    if (seg_len < 1) seg_len = 1;
    nseg = uint32_t((n + seg_len - 1) / seg_len);
    if (nseg == 0) nseg = 1;
}

struct SisiLaunch {
    si_context* ctx;
    uint32_t nseg;
    const uint32_t* base;
    uint64_t oA, nA, oB, nB, seg_len;
    uint64_t* partials;
    uint32_t* done;
    uint64_t* out;
    uint32_t* out32;
    uint32_t* stage;
    uint64_t cap;
    uint32_t* list;
    uint64_t* state;
    uint32_t ramp = 0;  // materialisation: segment length ramp (ramp_start)
};

template <int NT, int VT, int WN, bool SEL, bool MAT>
static void launch_stream(const SisiLaunch& L) {
    hipLaunchKernelGGL((k_sisi_stream<NT, VT, WN, MAT, SEL>), dim3(L.nseg), dim3(NT), 0, L.ctx->stream, L.base, L.oA,
                       L.nA, L.oB, L.nB, L.seg_len, L.nseg, L.partials, L.done, L.out, L.out32, L.stage, L.cap,
                       L.list, L.state, uint32_t(L.ctx->sisi_prio), MAT ? L.ramp : 0u);
}

template <bool MAT>
static void launch_shape(int shape, const SisiLaunch& L) {
    switch (shape) {
        case 1: launch_stream<256, 31, 4096, true, MAT>(L); break;
        case 2: launch_stream<128, 31, 2048, false, MAT>(L); break;
        case 3: launch_stream<256, 15, 2048, false, MAT>(L); break;
        case 4: launch_stream<128, 15, 1024, false, MAT>(L); break;
        case 5: launch_stream<512, 15, 4096, false, MAT>(L); break;
        case 6: launch_stream<256, 15, 4096, false, MAT>(L); break;
        default: launch_stream<256, 31, 4096, false, MAT>(L); break;
    }
}

// The default shape for a pair: 256 x 31 (T = 7936) takes about T/2 of each
// side per tile; when one set is at least 3x the other a tile would take more
// than W - 4 of the larger and be cut (threads idle), so 256 x 15 (T = 3840 <=
// W - 4) is used.  Measured at |A| = 2^27, |B| = 2^16: count 0.261 -> 0.198 ms.
static int auto_shape(uint64_t nA, uint64_t nB) {
    const uint64_t lo = nA < nB ? nA : nB, hi = nA < nB ? nB : nA;
    return hi >= 3 * lo ? 6 : 0;
}

// si_params.reserved[0]: 0 = automatic, else a shape index.  Counting alone
// takes the one-read-per-step variant of the 256 x 31 tile (shape 1): fewer
// LDS reads, measured 0.518 vs 0.539 ms at 2^28 x 2^28; the materialising
// kernel, which re-reads the window for its matches, is faster without it
// (0.566 vs 0.591 ms).
static int shape_of(const si_params* p, uint64_t nA, uint64_t nB) {
    const int s = p ? int(p->reserved[0]) : 0;
    if (s > 0 && s < kNumShapes) return s;
    const int a = auto_shape(nA, nB);
    return a == 0 ? 1 : a;
}

int ip_pair_count(si_dataset* ds, uint32_t a, uint32_t b, const si_params* p, uint64_t* d_out, uint32_t* d_out32) {
    si_context* ctx = ds->ctx;
    const uint64_t oA = ds->offsets[a], nA = ds->sizes[a];
    const uint64_t oB = ds->offsets[b], nB = ds->sizes[b];
    const int shape = shape_of(p, nA, nB);
    uint32_t nseg;
    uint64_t seg_len;
    sisi_geometry(ctx, nA + nB, p, kShapes[shape], nseg, seg_len);
    void* ws = nullptr;
    SI_TRY(small_workspace(ctx, 256 + uint64_t(nseg) * 8, &ws));
    if (nA == 0 || nB == 0) return zero_outputs(ctx, d_out, d_out32);
    SisiLaunch L{ctx, nseg, ds->d_elems, oA, nA, oB, nB, seg_len,
                 reinterpret_cast<uint64_t*>(static_cast<char*>(ws) + 256), static_cast<uint32_t*>(ws), d_out,
                 d_out32, nullptr, 0, nullptr, nullptr};
    KernelScope ks(ctx, "ip_sisi_count");
    launch_shape<false>(shape, L);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int ip_pair_materialize(si_dataset* ds, uint32_t a, uint32_t b, uint32_t* d_out, uint64_t* d_count) {
    si_context* ctx = ds->ctx;
    const uint64_t oA = ds->offsets[a], nA = ds->sizes[a];
    const uint64_t oB = ds->offsets[b], nB = ds->sizes[b];
    if (nA == 0 || nB == 0) {
        SI_HIP(hipMemsetAsync(d_count, 0, 8, ctx->stream));
        return SI_OK;
    }
    const int shape = ctx->sisi_shape > 0 && ctx->sisi_shape < kNumShapes ? ctx->sisi_shape : auto_shape(nA, nB);
    uint32_t nseg;
    uint64_t seg_len;
    sisi_geometry(ctx, nA + nB, nullptr, kShapes[shape], nseg, seg_len);
    // a segment of L merged elements holds at most L / 2 matches; with the
    // ramp the last segment is the longest
    const uint32_t ramp =
        ctx->sisi_ramp > 0 && ctx->sisi_ramp < 128 && nseg >= 2 && nseg < 4096 ? uint32_t(ctx->sisi_ramp) : 0u;
    const uint64_t maxLen = ramp ? ramp_start(nA + nB, nseg, ramp, nseg) - ramp_start(nA + nB, nseg, ramp, nseg - 1)
                                 : seg_len;
    const uint64_t cap = (((maxLen > seg_len ? maxLen : seg_len) / 2 + 2) + 63) & ~uint64_t(63);
    void* ws = nullptr;
    SI_TRY(small_workspace(ctx, 256, &ws));
    SI_TRY(ctx->ws_stage.ensure(cap * nseg * sizeof(uint32_t)));
    // look-back states: zeroed once when (re)allocated, left zeroed by every launch
    if (ctx->ws_lookback.bytes < uint64_t(nseg) * 8) {
        SI_TRY(ctx->ws_lookback.ensure(std::max<uint64_t>(uint64_t(nseg) * 8, 1 << 14)));
        SI_HIP(hipMemsetAsync(ctx->ws_lookback.ptr, 0, ctx->ws_lookback.bytes, ctx->stream));
    }
    SisiLaunch L{ctx, nseg, ds->d_elems, oA, nA, oB, nB, seg_len, nullptr, static_cast<uint32_t*>(ws), d_count,
                 nullptr, ctx->ws_stage.as<uint32_t>(), cap, d_out, ctx->ws_lookback.as<uint64_t>(), ramp};
    KernelScope ks(ctx, "ip_sisi_materialize");
    launch_shape<true>(shape, L);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int ip_run(si_dataset* ds, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts) {
    si_context* ctx = ds->ctx;
    const bool self = (A.id == B.id);
    const uint64_t aLen = A.end - A.start, bLen = B.end - B.start;
    const uint64_t npairs = self ? (aLen * (aLen - (aLen > 0)) / 2) : aLen * bLen;
    if (npairs == 0) return SI_OK;
    uint64_t grid = uint64_t(ctx->num_cus) * 32;
    if (p && p->blocks) grid = p->blocks;
    if (grid > npairs) grid = npairs;
    KernelScope ks(ctx, "ip_run");
    if (self)
        hipLaunchKernelGGL((k_run_merge<RUN_VT, true>), dim3(uint32_t(grid)), dim3(WAVE), 0, ctx->stream, ds->d_elems,
                           ds->d_offsets, A.start, uint32_t(aLen), A.id, B.start, uint32_t(bLen), B.id, n, npairs,
                           d_counts);
    else
        hipLaunchKernelGGL((k_run_merge<RUN_VT, false>), dim3(uint32_t(grid)), dim3(WAVE), 0, ctx->stream, ds->d_elems,
                           ds->d_offsets, A.start, uint32_t(aLen), A.id, B.start, uint32_t(bLen), B.id, n, npairs,
                           d_counts);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

}  // namespace si_impl
