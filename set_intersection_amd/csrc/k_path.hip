// Merge-path partitioned OBS and hash intersection: the --path variants of
// obs and hi for gfx950.
//
// Reference behaviour: with --path, findDiagonals (helpers.cuh:74-166) cuts
// every pair's merge path into one partition per block, and each block then
// intersects its partition with the technique: intersectPathOBS
// (obs.cu:438-522) caches blockDim samples of the partition's A part in
// shared memory and binary-searches each element of the B part in it;
// intersectPathHI (hi.cu:420-479) hashes the partition's A part into global
// buckets and probes them with the B part.  The reference's versions are
// correct only for a single pair: intersectPathOBS indexes A without its set
// offset (obs.cu:506) and intersectPathHI ignores the run's tiles
// (hi.cu:431-433); neither defect is kept.
//
// MI355X design:
//  * A segment is a contiguous piece of the merge path.  Under the A-first tie
//    order (A[x] precedes B[y] iff A[x] <= B[y]) the only B element that can
//    match an A element of a segment outside the segment's own B part is the
//    B element that follows it, so a segment intersects A[ia, x1) with
//    B[ib, y1] (one B element more); no diagonal buffer and no "count
//    augmentation" (helpers.cuh:121-123) are needed.
//  * SISI (k_path_pair): the merge path is cut into one contiguous range per
//    workgroup; a workgroup walks its range in segments of L merged elements,
//    each starting where the previous one ended, so only the range start
//    needs a full co-rank search and each segment end is searched inside a
//    window of L + 1 candidates (32-ary, ceil(log32 L) rounds).
//  * Runs (k_path_run): one wave per pair of the run, walking the pair's merge
//    path in segments the same way; counts are stored once per pair.
//  * Per segment, OBS caches NS samples of the larger part in LDS and searches
//    each element of the other part (phase 1 in LDS, phase 2 in global
//    memory); HASH inserts the smaller part into an LDS open-addressing table
//    (load <= 1/2) and probes it with the other part.  Both parts are read as
//    aligned 16-byte chunks.
#include "si_device.h"
#include "si_internal.h"

namespace {

using namespace si_dev;

enum { PATH_OBS = 0, PATH_HASH = 1 };
constexpr uint32_t EMPTY = 0xFFFFFFFFu;  // empty table slot; the value itself sets hasMax

// TS: LDS words of the table (HASH) or of the sample cache (OBS).
template <int NT, int TS>
struct PathSmem {
    alignas(16) uint32_t buf[TS];
    uint32_t hasMax;
    uint64_t red[NT / WAVE];
};

// Calls f(x) for every element x of base[e0, e1) (absolute indices) read as
// aligned 16-byte chunks: thread t takes chunks c0 + t + j*NT, four loads in
// flight.  Chunks past the range are clamped to its last chunk (in bounds,
// not used).  The element array is padded to a whole chunk.
template <int NT, typename F>
__device__ __forceinline__ void for_each_chunked(const uint32_t* __restrict__ base, uint64_t e0, uint64_t e1, F&& f) {
    if (e0 >= e1) return;
    const uint4* b4 = reinterpret_cast<const uint4*>(base);
    const uint64_t c0 = e0 >> 2, c1 = (e1 + 3) >> 2;
    constexpr int UNR = 4;
    for (uint64_t c = c0 + threadIdx.x; c < c1; c += uint64_t(UNR) * NT) {
        uint4 v[UNR];
#pragma unroll
        for (int j = 0; j < UNR; ++j) {
            const uint64_t cj = c + uint64_t(j) * NT;
            v[j] = b4[cj < c1 ? cj : c1 - 1];
        }
#pragma unroll
        for (int j = 0; j < UNR; ++j) {
            const uint64_t cj = c + uint64_t(j) * NT;
            const uint64_t e = cj << 2;
            const uint32_t x[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (cj < c1 && e + q >= e0 && e + q < e1) f(x[q]);
        }
    }
}

__device__ __forceinline__ uint32_t hslot(uint32_t x, uint32_t logTS) { return (x * 0x9E3779B1u) >> (32 - logTS); }

// Matches of one segment: |base[h0, h0+nh) ∩ base[q0, q0+nq)| where the first
// range is the part that is hashed (HASH) or sampled (OBS).  Contains
// barriers; the caller separates consecutive calls by one more barrier.
template <int NT, int TECH, int TS>
__device__ uint32_t segment_count(const uint32_t* __restrict__ base, uint64_t h0, uint64_t nh, uint64_t q0,
                                  uint64_t nq, PathSmem<NT, TS>& sm) {
    uint32_t cnt = 0;
    if (TECH == PATH_HASH) {
        // nh <= TS / 2: a table of the next power of two >= 2 nh (>= 64) slots
        const uint32_t need = uint32_t(2 * nh);
        uint32_t logTS = 32u - uint32_t(__builtin_clz(need > 1 ? need - 1 : 1u));
        logTS = logTS < 6 ? 6 : logTS;
        const uint32_t ts = 1u << logTS, mask = ts - 1;
        uint4* t4 = reinterpret_cast<uint4*>(sm.buf);
        for (uint32_t i = threadIdx.x; i < ts / 4; i += NT) t4[i] = make_uint4(EMPTY, EMPTY, EMPTY, EMPTY);
        if (threadIdx.x == 0) sm.hasMax = 0;
        __syncthreads();
        for_each_chunked<NT>(base, h0, h0 + nh, [&](uint32_t x) {
            if (x == EMPTY) {
                sm.hasMax = 1;
                return;
            }
            uint32_t s = hslot(x, logTS);
            while (true) {
                const uint32_t prev = atomicCAS(&sm.buf[s], EMPTY, x);
                if (prev == EMPTY || prev == x) return;
                s = (s + 1) & mask;
            }
        });
        __syncthreads();
        const uint32_t hm = sm.hasMax;
        for_each_chunked<NT>(base, q0, q0 + nq, [&](uint32_t x) {
            if (x == EMPTY) {
                cnt += hm;
                return;
            }
            uint32_t s = hslot(x, logTS);
            while (true) {
                const uint32_t k = sm.buf[s];
                if (k == x) {
                    ++cnt;
                    return;
                }
                if (k == EMPTY) return;
                s = (s + 1) & mask;
            }
        });
    } else {
        // samples S[t * nS / NS] of the larger part (t < NS = TS)
        const uint32_t* S = base + h0;
        for (uint32_t t = threadIdx.x; t < uint32_t(TS); t += NT) sm.buf[t] = S[(uint64_t(t) * nh) / TS];
        __syncthreads();
        for_each_chunked<NT>(base, q0, q0 + nq, [&](uint32_t x) { cnt += sample_search<TS>(S, nh, sm.buf, x); });
    }
    return cnt;
}

// The segment's two parts A[ia, x1) and B[ib, yb) (yb = y1 + 1, or nB at the
// end of B) in the roles of the technique: HASH hashes the smaller part, OBS
// samples the larger one.
template <int NT, int TECH, int TS>
__device__ __forceinline__ uint32_t segment(const uint32_t* __restrict__ base, uint64_t oA, uint64_t ia, uint64_t x1,
                                            uint64_t oB, uint64_t ib, uint64_t yb, PathSmem<NT, TS>& sm) {
    const uint64_t na = x1 - ia, nb = yb - ib;
    if (na == 0 || nb == 0) return 0;  // uniform over the workgroup
    const bool aFirst = (TECH == PATH_HASH) ? (na <= nb) : (na >= nb);
    return aFirst ? segment_count<NT, TECH, TS>(base, oA + ia, na, oB + ib, nb, sm)
                  : segment_count<NT, TECH, TS>(base, oB + ib, nb, oA + ia, na, sm);
}

// SISI: workgroup w intersects the merge-path range [w*range, (w+1)*range) of
// (A, B) in segments of L merged elements.  L = TS keeps the hashed part
// (at most L/2 elements) at load <= 1/2; OBS takes L = 8 TS.
template <int NT, int TECH, int TS>
__global__ __launch_bounds__(NT) void k_path_pair(const uint32_t* __restrict__ base, uint64_t oA, uint64_t nA,
                                                  uint64_t oB, uint64_t nB, uint64_t range,
                                                  uint64_t* __restrict__ partials, uint32_t* __restrict__ done,
                                                  uint64_t* __restrict__ out64, uint32_t* __restrict__ out32) {
    __shared__ PathSmem<NT, TS> sm;
    constexpr uint64_t L = TECH == PATH_HASH ? uint64_t(TS) : 8ull * TS;
    const uint64_t n = nA + nB;
    const uint64_t r0 = uint64_t(blockIdx.x) * range;
    const uint64_t rE = r0 + range < n ? r0 + range : n;
    const uint32_t* A = base + oA;
    const uint32_t* B = base + oB;
    uint64_t cnt = 0;
    if (r0 < rE) {
        // every wave finds the co-ranks itself (same loads, no broadcast)
        uint64_t ia = corank_halfwave(A, nA, B, nB, r0);
        uint64_t ib = r0 - ia;
        for (uint64_t d = r0; d < rE;) {
            const uint64_t d1 = d + L < rE ? d + L : rE;
            const uint64_t lo = d1 > nB ? (d1 - nB > ia ? d1 - nB : ia) : ia;
            const uint64_t hi = ia + (d1 - d) < nA ? ia + (d1 - d) : nA;
            const uint64_t x1 = corank_halfwave_range(A, B, d1, lo, hi);
            const uint64_t y1 = d1 - x1;
            cnt += segment<NT, TECH, TS>(base, oA, ia, x1, oB, ib, y1 < nB ? y1 + 1 : nB, sm);
            ia = x1;
            ib = y1;
            d = d1;
            __syncthreads();  // the segment's LDS reads are done before the next one rewrites it
        }
    }
    grid_fold_u64<NT>(cnt, partials, done, out64, out32, sm.red);
}

// Runs: one wave per pair (grid-stride), the pair's merge path walked in
// segments of L merged elements.
template <int TECH, int TS, bool SELF>
__global__ __launch_bounds__(WAVE) void k_path_run(const uint32_t* __restrict__ base,
                                                   const uint64_t* __restrict__ offsets, RunArgs r, uint64_t npairs,
                                                   uint32_t* __restrict__ counts) {
    __shared__ PathSmem<WAVE, TS> sm;
    constexpr uint64_t L = TECH == PATH_HASH ? uint64_t(TS) : 8ull * TS;
    const uint32_t aLen = r.aEnd - r.aStart, bLen = r.bEnd - r.bStart;
    for (uint64_t p = blockIdx.x; p < npairs; p += gridDim.x) {
        uint32_t li, lj;
        if (SELF) {
            tri_decode(p, aLen, li, lj);
        } else {
            li = uint32_t(p / bLen);
            lj = uint32_t(p % bLen);
        }
        const uint32_t a = r.aStart + li, b = r.bStart + lj;
        const uint64_t oA = offsets[a], nA = offsets[a + 1] - oA;
        const uint64_t oB = offsets[b], nB = offsets[b + 1] - oB;
        uint32_t cnt = 0;
        if (nA && nB) {
            const uint64_t n = nA + nB;
            uint64_t ia = 0, ib = 0;
            for (uint64_t d = 0; d < n;) {
                const uint64_t d1 = d + L < n ? d + L : n;
                uint64_t x1 = nA;
                if (d1 < n) {
                    const uint64_t lo = d1 > nB ? (d1 - nB > ia ? d1 - nB : ia) : ia;
                    const uint64_t hi = ia + (d1 - d) < nA ? ia + (d1 - d) : nA;
                    x1 = corank_halfwave_range(base + oA, base + oB, d1, lo, hi);
                }
                const uint64_t y1 = d1 - x1;
                cnt += segment<WAVE, TECH, TS>(base, oA, ia, x1, oB, ib, y1 < nB ? y1 + 1 : nB, sm);
                ia = x1;
                ib = y1;
                d = d1;
                __syncthreads();
            }
        }
        cnt = wave_sum_u32(cnt);
        if (lane_id() == 0) counts[run_index<SELF>(r, a, b)] = cnt;
    }
}

constexpr int PAIR_NT = 256;
constexpr int PAIR_TS_HASH = 8192;  // 32 KiB table, segments of 8192 merged elements
constexpr int PAIR_NS_OBS = 256;    // samples per segment (the reference caches blockDim samples)
constexpr int RUN_TS_HASH = 2048;   // 8 KiB per wave
constexpr int RUN_NS_OBS = 64;

}  // namespace

namespace si_impl {

// SISI --path for obs (tech == SI_TECH_OBS) and hi (SI_TECH_HI).
int path_pair_count(si_dataset* ds, uint32_t a, uint32_t b, int32_t tech, const si_params* p, uint64_t* d_out,
                    uint32_t* d_out32) {
    si_context* ctx = ds->ctx;
    const uint64_t oA = ds->offsets[a], nA = ds->sizes[a], oB = ds->offsets[b], nB = ds->sizes[b];
    if (nA == 0 || nB == 0) return zero_outputs(ctx, d_out, d_out32);
    const uint64_t n = nA + nB;
    const bool hash = tech == SI_TECH_HI;
    // one contiguous range of the merge path per workgroup, at least one
    // segment long; four workgroups per CU (the hash table is 32 KiB of LDS)
    const uint64_t L = hash ? uint64_t(PAIR_TS_HASH) : 8ull * PAIR_NS_OBS;
    uint64_t g = (p && p->blocks) ? p->blocks : uint64_t(ctx->num_cus) * 4;
    g = std::max<uint64_t>(1, std::min<uint64_t>(g, (n + L - 1) / L));
    const uint64_t range = (n + g - 1) / g;
    g = (n + range - 1) / range;
    void* ws = nullptr;
    SI_TRY(small_workspace(ctx, 256 + g * 8, &ws));
    uint64_t* partials = reinterpret_cast<uint64_t*>(static_cast<char*>(ws) + 256);
    uint32_t* done = static_cast<uint32_t*>(ws);
    if (hash) {
        KernelScope ks(ctx, "hi_path_pair");
        hipLaunchKernelGGL((k_path_pair<PAIR_NT, PATH_HASH, PAIR_TS_HASH>), dim3(uint32_t(g)), dim3(PAIR_NT), 0,
                           ctx->stream, ds->d_elems, oA, nA, oB, nB, range, partials, done, d_out, d_out32);
    } else {
        KernelScope ks(ctx, "obs_path_pair");
        hipLaunchKernelGGL((k_path_pair<PAIR_NT, PATH_OBS, PAIR_NS_OBS>), dim3(uint32_t(g)), dim3(PAIR_NT), 0,
                           ctx->stream, ds->d_elems, oA, nA, oB, nB, range, partials, done, d_out, d_out32);
    }
    SI_HIP(hipGetLastError());
    return SI_OK;
}

// Runs with --path for obs and hi: every pair of the run through the
// merge-path segments of one wave.
int path_run(si_dataset* ds, int32_t tech, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts) {
    si_context* ctx = ds->ctx;
    const bool self = (A.id == B.id);
    const uint64_t aLen = A.end - A.start, bLen = B.end - B.start;
    const uint64_t npairs = self ? (aLen * (aLen - (aLen > 0)) / 2) : aLen * bLen;
    if (npairs == 0) return SI_OK;
    RunArgs r{A.start, A.end, A.id, B.start, B.end, B.id, n};
    uint64_t grid = (p && p->blocks) ? p->blocks : uint64_t(ctx->num_cus) * 16;
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, npairs));
    const bool hash = tech == SI_TECH_HI;
    KernelScope ks(ctx, hash ? "hi_path_run" : "obs_path_run");
#define LAUNCH(T, TS, S)                                                                                       \
    hipLaunchKernelGGL((k_path_run<T, TS, S>), dim3(uint32_t(grid)), dim3(WAVE), 0, ctx->stream, ds->d_elems, \
                       ds->d_offsets, r, npairs, d_counts)
    if (hash) {
        if (self) LAUNCH(PATH_HASH, RUN_TS_HASH, true); else LAUNCH(PATH_HASH, RUN_TS_HASH, false);
    } else {
        if (self) LAUNCH(PATH_OBS, RUN_NS_OBS, true); else LAUNCH(PATH_OBS, RUN_NS_OBS, false);
    }
#undef LAUNCH
    SI_HIP(hipGetLastError());
    return SI_OK;
}

}  // namespace si_impl
