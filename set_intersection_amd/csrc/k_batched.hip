// Batched all-pairs (MISI) dispatcher for gfx950.
//
// Replaces the reference's per-run loops (ip.cu:141-177; the same loop in
// obs.cu:173-243, hi.cu:198-261, bitmap_dynamic.cu:363-389), which give every
// pair of a run the same kernel and re-read both sets for every pair.
//
// Design:
//  * Plan (once per dataset): sets are ordered by size, descending, and the
//    elements are gathered into that order on the device.  Every pair (i, j),
//    i < j in that order, is charged to its larger set i (the "row"); the
//    partners of row i are the sets i+1 .. k-1, whose elements are ONE
//    contiguous range of the reordered element array.
//  * Bins: a row is BITMAP (|S_i| >= bitmap_min and the universe fits a
//    2^20-bit LDS bitmap), HASH (an LDS open-addressing table of 2|S_i|
//    slots fits) or BIG.  For a BIG row the partners with
//    |S_i| >= ratio_search * |S_j| are SEARCH partners (binary search in S_i,
//    first level from an LDS sample cache) and the rest are MERGE pairs (merge
//    path).  So each pair costs about min(|S_i|, |S_j|) LDS probes instead of
//    |S_i| + |S_j| merge steps.
//  * Work items are (row, partner-set range) cut at set boundaries into
//    ~chunk_elems partner elements.  Persistent workgroups pop items from an
//    atomic queue; a workgroup rebuilds its row structure only when the row
//    changes.  Inside an item the waves split the partner elements evenly,
//    stream them with coalesced loads, probe, and turn the hit bits into
//    per-pair counts with wave ballots + popcount over each set's bit range;
//    a set cut by a wave boundary is completed through LDS.  Every count is
//    written exactly once with a plain store, so the output needs no clearing
//    and the launch sequence (queue reset + kernels) can be captured in a
//    hipGraph and replayed.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "si_device.h"
#include "si_internal.h"
#include "si_merge.h"

namespace {

using namespace si_dev;

enum : uint32_t { MODE_BITMAP = 0, MODE_HASH = 1, MODE_SEARCH = 2 };

struct Item {
    uint32_t row, j0, j1, mode;
};

constexpr uint32_t EMPTY = 0xFFFFFFFFu;
constexpr int BM_NT = 1024, BM_WORDS = 32768;  // 128 KiB LDS bitmap: universes up to 2^20
constexpr int HB_WORDS = 16384;                // 64 KiB LDS bitmap: a universe half, two workgroups per CU
constexpr int PR_NT = 256, PR_WORDS = 8192;    // 32 KiB LDS: hash tables of <= 8192 slots
constexpr uint32_t HASH_MAX = PR_WORDS / 2;    // largest row that fits a hash table
constexpr int SEARCH_NS = 4096;                // samples cached for SEARCH rows
constexpr int VT = 24;                         // elements per lane per chunk (24 loads in flight: 27.6 -> 24.7 ms for MISI 16384 Zipf)
// zeroed elements after the reordered element array: a wave's whole chunk
// (WAVE * VT positions) starting at any element stays inside the allocation
constexpr uint32_t SELEM_PAD = WAVE * VT;

__device__ __forceinline__ uint32_t hslot(uint32_t x, uint32_t logTS) { return (x * 0x9E3779B1u) >> (32 - logTS); }

// tri_index(k, lo, hi) with 32-bit factors: one 32 x 32 -> 64-bit multiply
// per lane (k(k-1)/2 is wave-uniform)
__device__ __forceinline__ uint64_t pair_slot(uint32_t k, uint32_t pa, uint32_t pb) {
    const uint32_t lo = pa < pb ? pa : pb, hi = pa < pb ? pb : pa;
    const uint32_t m = k - lo;  // >= 2: lo < hi < k
    return (uint64_t(k) * (k - 1) >> 1) - (uint64_t(m) * (m - 1) >> 1) + (hi - lo - 1);
}

// First index j in [lo, hi) with pass(j), or hi if none (pass monotone:
// false ... false true ... true), found by the whole wave: each round tests
// 64 indices spread over the candidates, so n candidates take ceil(log64 n)
// dependent rounds instead of the log2 n of a binary search.  All lanes call
// it with the same arguments.
template <typename Pass>
__device__ __forceinline__ uint32_t wave_first(uint32_t lo, uint32_t hi, Pass pass) {
    const uint32_t lane = lane_id();
    while (lo < hi) {  // the answer lies in [lo, hi]
        const uint32_t n = hi - lo, step = (n + WAVE - 1) / WAVE;
        const uint32_t idx = lo + lane * step;
        const uint64_t b = __ballot(idx < hi && pass(idx));
        if (b) {  // the first passing lane bounds the answer from above
            const uint32_t f = uint32_t(__builtin_ctzll(b));
            hi = lo + f * step;
            lo = f ? lo + (f - 1) * step + 1 : lo;
        } else {  // every tested index fails: the answer follows the last one
            lo += ((n - 1) / step) * step + 1;
        }
    }
    return lo;
}

// v_writelane_b32: lane L of `old` replaced by the wave-uniform value `val`
// (no compare, no select)
template <int L>
__device__ __forceinline__ uint32_t write_lane(uint32_t old, uint32_t val) {
    uint32_t r;
    asm("v_writelane_b32 %0, %1, %2" : "=v"(r) : "s"(val), "i"(L), "0"(old));
    return r;
}

// words V.. of a chunk's masks into their lanes (compile-time lane numbers)
template <int V>
__device__ __forceinline__ void lane_masks(const uint64_t (&M)[VT], uint32_t& mlo, uint32_t& mhi) {
    if constexpr (V < VT) {
        mlo = write_lane<V>(mlo, uint32_t(M[V]));
        mhi = write_lane<V>(mhi, uint32_t(M[V] >> 32));
        lane_masks<V + 1>(M, mlo, mhi);
    }
}

// x + (x of lane l - d within the lane's row of 16; 0 for the row's first d lanes)
template <int D>
__device__ __forceinline__ uint32_t add_row_shr(uint32_t x) {
    return x + uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x110 + D, 0xf, 0xf, false));
}

// A wave's chunk of 64 x VT positions is probed into VT ballot masks (bit l
// of word v: position 64 v + l).  The wave stores the masks and the hits
// before each word (entry VT: no bits, the chunk's total) in its own LDS
// arrays (VT + 1 entries each), so that the hits at chunk positions [a, b)
// are chunk_hits_before(b) - chunk_hits_before(a): two LDS reads and two
// popcounts per lane instead of masking all VT words.
__device__ __forceinline__ void store_chunk_masks(const uint64_t (&M)[VT], uint64_t* wm, uint32_t* wp) {
    static_assert(VT + 1 <= 32, "the prefix scan covers lanes 0..31");
    const uint32_t lane = lane_id();
    // lane v takes word v (one v_writelane per half: the masks are
    // wave-uniform); the hits before word v are an exclusive prefix sum of
    // the words' popcounts over lanes 0..VT (lanes >= VT hold no bits):
    // a scan within rows of 16 lanes, then row 0's total added to row 1
    uint32_t mlo = 0, mhi = 0;
    lane_masks<0>(M, mlo, mhi);
    const uint32_t c = uint32_t(__builtin_popcount(mlo)) + uint32_t(__builtin_popcount(mhi));
    uint32_t x = add_row_shr<1>(c);
    x = add_row_shr<2>(x);
    x = add_row_shr<4>(x);
    x = add_row_shr<8>(x);
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x142, 0xa, 0xf, false));  // row_bcast:15
    if (lane <= uint32_t(VT)) {
        wm[lane] = (uint64_t(mhi) << 32) | mlo;
        wp[lane] = x - c;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// lanes l of word v whose chunk position 64 v + l is at most `last`
// (wave-uniform: scalar instructions)
__device__ __forceinline__ uint64_t lanes_upto(uint32_t last, int v) {
    const uint32_t b = uint32_t(v) * WAVE;
    return last >= b + (WAVE - 1) ? ~0ull : last < b ? 0ull : (2ull << (last - b)) - 1ull;
}

__device__ __forceinline__ uint32_t chunk_hits_before(const uint64_t* wm, const uint32_t* wp, uint32_t x) {
    const uint32_t v = x >> 6, r = x & 63u;  // x in [0, 64 VT]
    return wp[v] + uint32_t(__popcll(wm[v] & ((1ull << r) - 1ull)));
}

template <int LDSW>
struct RowSmem {
    alignas(16) uint32_t tab[LDSW];
    uint32_t item, row, mode, hasMax, logTS;
    uint32_t pset[2 * 32], pcnt[2 * 32];
};

// BITMAP rows: bit x - base for each element x (all in [base, base + 32 * bmWords)).
template <int NT, int LDSW>
__device__ void build_row(RowSmem<LDSW>& sm, const uint32_t* __restrict__ S, uint64_t n, uint32_t mode,
                          uint32_t bmWords, uint32_t base = 0) {
    if (mode == MODE_BITMAP) {
        // 16-byte clears, then the sorted row's bits by wave chunks with eight
        // chunk loads in flight (plain stores inside a chunk, atomics at its
        // edges)
        for (uint32_t i = threadIdx.x; i < bmWords / 4; i += NT)
            reinterpret_cast<uint4*>(sm.tab)[i] = make_uint4(0u, 0u, 0u, 0u);
        for (uint32_t i = (bmWords & ~3u) + threadIdx.x; i < bmWords; i += NT) sm.tab[i] = 0;
        __syncthreads();
        constexpr uint32_t NW = NT / WAVE;
        const uint64_t chunks = (n + WAVE - 1) / WAVE;
        for (uint64_t c = threadIdx.x / WAVE; c < chunks; c += uint64_t(NW) * BUILD_UNROLL)
            or_bits_chunks(S, n, c, NW, sm.tab, base);
    } else if (mode == MODE_HASH) {
        uint32_t logTS = 6;
        // load <= 1/16 where the LDS region allows (short probe chains), else <= 1/2
        while ((1ull << logTS) < 16 * n && (1ull << logTS) < uint64_t(LDSW)) ++logTS;
        while ((1ull << logTS) < 2 * n) ++logTS;
        const uint32_t TS = 1u << logTS;
        for (uint32_t i = threadIdx.x; i < TS; i += NT) sm.tab[i] = EMPTY;
        if (threadIdx.x == 0) {
            sm.hasMax = 0;
            sm.logTS = logTS;
        }
        __syncthreads();
        for (uint64_t i = threadIdx.x; i < n; i += NT) {
            const uint32_t x = S[i];
            if (x == EMPTY) {
                sm.hasMax = 1;
                continue;
            }
            uint32_t s = hslot(x, logTS);
            while (true) {
                const uint32_t prev = atomicCAS(&sm.tab[s], EMPTY, x);
                if (prev == EMPTY || prev == x) break;
                s = (s + 1) & (TS - 1);
            }
        }
    } else {  // SEARCH: sample cache
        for (uint32_t t = threadIdx.x; t < SEARCH_NS; t += NT) sm.tab[t] = S[(uint64_t(t) * n) / SEARCH_NS];
    }
    __syncthreads();
}

template <int LDSW>
__device__ __forceinline__ uint32_t probe(const RowSmem<LDSW>& sm, uint32_t mode, uint32_t x,
                                          const uint32_t* __restrict__ S, uint64_t nS, uint32_t logTS, uint32_t hasMax) {
    if (mode == MODE_BITMAP) return (sm.tab[x >> 5] >> (x & 31)) & 1u;
    if (mode == MODE_HASH) {
        if (x == EMPTY) return hasMax;
        const uint32_t mask = (1u << logTS) - 1;
        uint32_t s = hslot(x, logTS);
        while (true) {
            const uint32_t k = sm.tab[s];
            if (k == x) return 1;
            if (k == EMPTY) return 0;
            s = (s + 1) & mask;
        }
    }
    // SEARCH
    uint32_t lo = 0, hi = SEARCH_NS;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sm.tab[mid] <= x) lo = mid + 1; else hi = mid;
    }
    if (lo == 0) return 0;
    if (sm.tab[lo - 1] == x) return 1;
    uint64_t l = (uint64_t(lo - 1) * nS) / SEARCH_NS + 1;
    uint64_t h = (lo == SEARCH_NS) ? nS : (uint64_t(lo) * nS) / SEARCH_NS;
    while (l < h) {
        const uint64_t mid = (l + h) >> 1;
        const uint32_t y = S[mid];
        if (y == x) return 1;
        if (y < x) l = mid + 1; else h = mid;
    }
    return 0;
}

// Persistent row kernel: pops items from queue[0], builds the row's LDS
// structure, streams the item's partner elements and stores one count per pair.
template <int NT, int LDSW>
__global__ __launch_bounds__(NT) void k_rows(const Item* __restrict__ items, const uint32_t* __restrict__ starts,
                                             uint32_t nitems, uint32_t* __restrict__ queue,
                                             const uint32_t* __restrict__ selem, const uint64_t* __restrict__ soff,
                                             const uint32_t* __restrict__ perm, uint32_t k, uint32_t bmWords,
                                             uint32_t* __restrict__ out) {
    constexpr int NW = NT / WAVE;
    static_assert(NW <= 32, "partial slots");
    __shared__ RowSmem<LDSW> sm;
    __shared__ uint64_t wmask[NW][VT + 1];  // store_chunk_masks, per wave
    __shared__ uint32_t wpre[NW][VT + 1];
    // the wave number is wave-uniform: held in a scalar register, so the
    // wave's partner range and chunk bases are scalar too
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = __builtin_amdgcn_readfirstlane(tid / WAVE);
    if (tid == 0) {
        sm.row = EMPTY;
        sm.mode = EMPTY;
    }
    __syncthreads();
    while (true) {
        if (tid == 0) sm.item = atomicAdd(queue, 1u);
        if (tid < 2 * NW) sm.pset[tid] = EMPTY;
        __syncthreads();
        const uint32_t it = sm.item;
        if (it >= nitems) break;
        const Item I = items[it];
        // this wave's first partner set: the largest j in [j0, j1) with
        // soff[j] <= the wave's first partner element (from the plan)
        const uint32_t jcFirst = starts[uint64_t(it) * NW + w];
        if (I.row != sm.row || I.mode != sm.mode) {
            __syncthreads();  // everyone has read sm.row / sm.mode
            build_row<NT, LDSW>(sm, selem + soff[I.row], soff[I.row + 1] - soff[I.row], I.mode, bmWords);
            if (tid == 0) {
                sm.row = I.row;
                sm.mode = I.mode;
            }
        }
        const uint32_t mode = I.mode, logTS = sm.logTS, hasMax = sm.hasMax;
        const uint64_t rowOff = soff[I.row], rowN = soff[I.row + 1] - rowOff;
        const uint32_t* S = selem + rowOff;
        const uint32_t prow = perm[I.row];

        // this wave's share of the item's partner elements
        const uint64_t e0 = soff[I.j0], e1 = soff[I.j1], len = e1 - e0;
        const uint64_t ws = e0 + (len * w) / NW, we = e0 + (len * (w + 1)) / NW;
        if (ws < we) {
            uint32_t jc = jcFirst;
            uint32_t carry = 0;
            // bitmap rows load a chunk's elements while the previous chunk's
            // sets are counted: a scalar chunk base plus a lane byte offset
            // and an immediate per word (instruction offsets reach 4095
            // bytes: words 16.. use a second base); positions past the
            // wave's end read other sets' elements or the zeroed padding,
            // all inside the universe and so inside the bitmap
            const uint32_t lane4 = 4u * lane;
            uint32_t xb[VT];
            auto load_bitmap_chunk = [&](uint64_t c) {
                const char* cb0 = reinterpret_cast<const char*>(selem + c);
                const char* cb1 = cb0 + 16 * 4 * WAVE;
#pragma unroll
                for (int v = 0; v < VT; ++v)
                    xb[v] = *reinterpret_cast<const uint32_t*>((v < 16 ? cb0 + v * 4 * WAVE : cb1 + (v - 16) * 4 * WAVE) +
                                                               lane4);
            };
            if (mode == MODE_BITMAP) load_bitmap_chunk(ws);
            for (uint64_t c0 = ws; c0 < we; c0 += WAVE * VT) {
                const uint64_t c1 = (c0 + WAVE * VT < we) ? c0 + WAVE * VT : we;
                uint64_t M[VT];
                const uint32_t* cb = selem + c0;
                const uint32_t last = uint32_t(c1 - c0) - 1;
                // bounds and output numbers of the sets jc + t, t < 2 x 64,
                // loaded beside the chunk (sets past j1 read set j1 - 1's)
                uint64_t pBeg[2], pEnd[2];
                uint32_t pPerm[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const uint32_t j = jc + uint32_t(h) * WAVE + lane, jj = j < I.j1 ? j : I.j1 - 1;
                    pBeg[h] = soff[jj];
                    pEnd[h] = soff[jj + 1];
                    pPerm[h] = perm[jj];
                }
                // the chunk's counts from its masks: called at the end of each
                // mode's branch, so the waits in the counting are those of its
                // own path
                auto count_sets = [&](const uint64_t (&M)[VT]) {
                    store_chunk_masks(M, wmask[w], wpre[w]);
                    // sets overlapping [c0, c1): jc, jc+1, ...; lane t of group g
                    // takes jc + g + t.  A group returns whether the sets go on
                    // past it.  The first two groups use the bounds loaded beside
                    // the chunk (straight-line code, so their uses wait only for
                    // those loads, not for the next chunk's, which are younger);
                    // further groups load their own.
                    uint32_t nextJc = jc, nextCarry = 0;
                    auto group = [&](uint32_t g, uint64_t sBeg, uint64_t sEnd, uint32_t pj) -> bool {
                        const uint32_t j = jc + g + lane;
                        const bool valid = j < I.j1 && sBeg < c1;
                        const uint64_t valid_mask = __ballot(valid);
                        if (!valid_mask) return false;
                        uint32_t cnt = 0;
                        bool endsHere = false;
                        if (valid) {
                            const uint32_t a = uint32_t((sBeg > c0 ? sBeg : c0) - c0);
                            const uint32_t b = uint32_t((sEnd < c1 ? sEnd : c1) - c0);
                            cnt = chunk_hits_before(wmask[w], wpre[w], b) - chunk_hits_before(wmask[w], wpre[w], a);
                            if (g == 0 && lane == 0) cnt += carry;
                            endsHere = sEnd <= c1;
                            if (endsHere) {
                                if (sEnd <= we && sBeg >= ws) {
                                    out[pair_slot(k, prow, pj)] = cnt;
                                } else if (sEnd <= we) {
                                    sm.pset[2 * w] = j;  // started before this wave
                                    sm.pcnt[2 * w] = cnt;
                                }
                            } else if (c1 == we) {
                                sm.pset[2 * w + 1] = j;  // continues past this wave
                                sm.pcnt[2 * w + 1] = cnt;
                            }
                        }
                        // the last valid lane's set decides where the next chunk starts
                        const int last = 63 - __builtin_clzll(valid_mask);
                        const uint32_t lastJ = jc + g + uint32_t(last);
                        const uint32_t lastEnds = __builtin_amdgcn_readlane(uint32_t(endsHere), last);
                        const uint32_t lastCnt = __builtin_amdgcn_readlane(cnt, last);
                        if (lastEnds) {
                            nextJc = lastJ + 1;
                            nextCarry = 0;
                        } else {
                            nextJc = lastJ;
                            nextCarry = lastCnt;
                        }
                        return last == 63;
                    };
                    if (group(0, pBeg[0], pEnd[0], pPerm[0]) && group(WAVE, pBeg[1], pEnd[1], pPerm[1])) {
                        for (uint32_t g = 2 * WAVE;; g += WAVE) {
                            const uint32_t j = jc + g + lane, jj = j < I.j1 ? j : I.j1 - 1;
                            if (!group(g, soff[jj], soff[jj + 1], perm[jj])) break;
                        }
                    }
                    jc = nextJc;
                    carry = nextCarry;
                };
                if (mode == MODE_BITMAP) {
                    // the common case, branch-free: one bit-field extract and
                    // compare per position; a partial chunk's end is a scalar
                    // lane mask per word
                    const bool whole = last == uint32_t(WAVE * VT - 1);
                    uint32_t wd[VT];
#pragma unroll
                    for (int v = 0; v < VT; ++v) wd[v] = sm.tab[xb[v] >> 5];
#pragma unroll
                    for (int v = 0; v < VT; ++v)
                        M[v] = __builtin_amdgcn_ballot_w64(__builtin_amdgcn_ubfe(wd[v], xb[v], 1u) != 0u);
                    if (!whole) {
#pragma unroll
                        for (int v = 0; v < VT; ++v) M[v] &= lanes_upto(last, v);
                    }
                    // unconditional, so the counting below sees the same loads in
                    // flight on every path (a conditional issue makes every wait
                    // in it a wait for these loads too); the wave's last chunk
                    // re-reads its own positions, which are cached
                    load_bitmap_chunk(c1 < we ? c1 : c0);
                    count_sets(M);
                } else if (mode == MODE_HASH) {
                    // positions past c1 re-read c1 - 1 (masked off)
                    uint32_t x[VT];
#pragma unroll
                    for (int v = 0; v < VT; ++v) {
                        const uint32_t q = uint32_t(v) * WAVE + lane;
                        x[v] = cb[q < last ? q : last];
                    }
#pragma unroll
                    for (int v = 0; v < VT; ++v)
                        M[v] = __ballot(uint32_t(v) * WAVE + lane <= last &&
                                        probe<LDSW>(sm, MODE_HASH, x[v], S, rowN, logTS, hasMax));
                    count_sets(M);
                } else {
#pragma unroll
                    for (int v = 0; v < VT; ++v) {
                        const uint64_t p = c0 + uint64_t(v) * WAVE + lane;
                        uint32_t hit = 0;
                        if (p < c1) hit = probe<LDSW>(sm, mode, selem[p], S, rowN, logTS, hasMax);
                        M[v] = __ballot(hit);
                    }
                    count_sets(M);
                }
            }
        }
        __syncthreads();
        // complete the sets cut by wave boundaries
        if (tid == 0) {
            for (int s = 0; s < 2 * NW; ++s) {
                const uint32_t j = sm.pset[s];
                if (j == EMPTY) continue;
                uint32_t tot = 0;
                for (int t = s; t < 2 * NW; ++t)
                    if (sm.pset[t] == j) {
                        tot += sm.pcnt[t];
                        if (t != s) sm.pset[t] = EMPTY;
                    }
                out[pair_slot(k, prow, perm[j])] = tot;
            }
        }
        __syncthreads();
    }
}

// MERGE pairs by the streaming merge path, one workgroup per pair:
// pairs[p] = (set a, set b, output slot low, output slot high).
template <int NT, int MVT>
__global__ __launch_bounds__(NT) void k_merge_pairs(const uint4* __restrict__ pairs, uint32_t npairs,
                                                    const uint32_t* __restrict__ elems,
                                                    const uint64_t* __restrict__ offsets, uint32_t* __restrict__ out) {
    __shared__ MergeSmem<NT, MVT> sm;
    for (uint32_t p = blockIdx.x; p < npairs; p += gridDim.x) {
        const uint4 pr = pairs[p];
        const uint64_t oA = offsets[pr.x], nA = offsets[pr.x + 1] - oA, oB = offsets[pr.y], nB = offsets[pr.y + 1] - oB;
        uint64_t c = 0;
        if (nA && nB) c = merge_segment<NT, MVT, false>(elems, oA, nA, oB, nB, 0, 0, nA, nB, sm, nullptr);
        c = block_sum_u64<NT>(c, sm.red);
        if (threadIdx.x == 0) out[(uint64_t(pr.w) << 32) | pr.z] = uint32_t(c);
        __syncthreads();
    }
}

// Zero counts for every pair whose smaller set (in sorted order) is empty.
__global__ void k_zero_pairs(const uint32_t* __restrict__ perm, uint32_t k, uint32_t kne, uint32_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x; i < k; i += gridDim.x) {
        const uint32_t j0 = (i + 1 > kne) ? i + 1 : kne;
        const uint32_t pi = perm[i];
        for (uint32_t j = j0 + threadIdx.x; j < k; j += blockDim.x) out[pair_slot(k, pi, perm[j])] = 0;
    }
}

__global__ void k_gather(const uint32_t* __restrict__ elems, const uint64_t* __restrict__ offsets,
                         const uint32_t* __restrict__ perm, const uint64_t* __restrict__ soff, uint32_t k,
                         uint32_t* __restrict__ selem) {
    for (uint32_t s = blockIdx.x; s < k; s += gridDim.x) {
        const uint64_t src = offsets[perm[s]], dst = soff[s], n = soff[s + 1] - dst;
        for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) selem[dst + i] = elems[src + i];
    }
}


// ------------------------------------------------------------ edge lists --
// A partner record is uint4 {element offset low, high, inclusive end of the
// partner's elements within its item, output index}.  Items are (row,
// partner range [j0, j1), mode); the item's partner elements are numbered
// 0 .. end[j1-1] in partner order (a virtual concatenation).

// First partner record q in [j0, j1) whose elements end past position ws
// (part[q].z > ws; part[j1 - 1].z > ws holds).  Partners have at least one
// element, so ws = 0 needs no search.
__device__ __forceinline__ uint32_t first_partner(const uint4* __restrict__ part, uint32_t j0, uint32_t j1,
                                                  uint32_t ws) {
    if (ws == 0) return j0;
    return wave_first(j0, j1, [&](uint32_t q) { return part[q].z > ws; });
}

// Lanes hold 64 consecutive partners of a wave's element range; every
// element position finds its partner by a shuffle binary search over the
// lanes' start positions, is probed against the row structure, and each lane
// sums the hit bits of its own partner's positions (ballot + popcount).  A
// partner wholly inside [ws, we) gets its count stored; one cut by ws or we
// leaves its part in cutSet/cutCnt (slot 0: began before ws, slot 1: goes on
// past we).  Partners must have at least one element.
template <typename ProbeFn>
__device__ __forceinline__ void edge_batches(const uint4* __restrict__ part, uint32_t j0, uint32_t j1, uint32_t ws,
                                             uint32_t we, const uint32_t* __restrict__ elems, ProbeFn probeFn,
                                             uint32_t* __restrict__ out, uint32_t* cutSet, uint32_t* cutCnt,
                                             uint64_t* wm, uint32_t* wp) {
    const uint32_t lane = lane_id();
    const uint32_t lo = first_partner(part, j0, j1, ws);
    uint32_t qc = lo, carry = 0;
    uint32_t c0 = ws;
    while (c0 < we) {
        const uint32_t q = qc + lane;
        const bool valid = q < j1;
        const uint4 P = valid ? part[q] : make_uint4(0, 0, 0xFFFFFFFFu, 0);
        const uint32_t end = P.z;
        uint32_t start = __shfl_up(end, 1, WAVE);
        const uint32_t start0 = (qc == j0) ? 0u : part[qc - 1].z;
        if (lane == 0) start = start0;
        const uint32_t e63 = __shfl(end, WAVE - 1, WAVE);
        uint32_t c1 = c0 + WAVE * VT;
        c1 = c1 < we ? c1 : we;
        c1 = c1 < e63 ? c1 : e63;
        // partner of each position: partners have elements, so their starts
        // strictly increase; the partners starting inside a 64-position step
        // (usually one or two) are walked with scalar reads, and each lane
        // keeps the last one starting at or before its position
        uint32_t curS = uint32_t(__builtin_amdgcn_readfirstlane(int(start)));
        uint32_t curLo = uint32_t(__builtin_amdgcn_readfirstlane(int(P.x)));
        uint32_t curHi = uint32_t(__builtin_amdgcn_readfirstlane(int(P.y)));
        uint64_t M[VT];
#pragma unroll
        for (int v = 0; v < VT; ++v) {
            const uint32_t base = c0 + uint32_t(v) * WAVE, pos = base + lane;
            uint32_t sr = curS, lo = curLo, hi = curHi;
            uint64_t m = __ballot(valid && start >= base && start < base + WAVE && start > c0);
            while (m) {
                const int r = __builtin_ctzll(m);
                m &= m - 1;
                const uint32_t s_r = uint32_t(__builtin_amdgcn_readlane(int(start), r));
                const uint32_t lo_r = uint32_t(__builtin_amdgcn_readlane(int(P.x), r));
                const uint32_t hi_r = uint32_t(__builtin_amdgcn_readlane(int(P.y), r));
                if (pos >= s_r) {
                    sr = s_r;
                    lo = lo_r;
                    hi = hi_r;
                }
                curS = s_r;
                curLo = lo_r;
                curHi = hi_r;
            }
            const uint64_t orr = (uint64_t(hi) << 32) | lo;
            uint32_t hit = 0;
            if (pos < c1) hit = probeFn(elems[orr + (pos - sr)]);
            M[v] = __ballot(hit);
        }
        // this lane's partner covers [start, end); its positions in [c0, c1)
        store_chunk_masks(M, wm, wp);
        uint32_t cnt = 0;
        const bool touches = valid && start < c1 && end > c0;
        if (touches) {
            const uint32_t a = (start > c0 ? start : c0) - c0;
            const uint32_t b = (end < c1 ? end : c1) - c0;
            cnt = chunk_hits_before(wm, wp, b) - chunk_hits_before(wm, wp, a);
            if (lane == 0) cnt += carry;
            if (end <= c1) {
                if (start >= ws && end <= we) out[P.w] = cnt;                  // whole partner in this wave
                else if (cutSet) { cutSet[0] = q; cutCnt[0] = cnt; }         // began before ws
            } else if (c1 == we && cutSet) {
                cutSet[1] = q;                                                // continues past we
                cutCnt[1] = cnt;
            }
        }
        // next batch starts at the first partner not finished by c1
        const uint64_t done = __ballot(valid && end <= c1);
        const uint32_t nd = uint32_t(__popcll(done));
        const uint32_t contCnt = __shfl(cnt, nd < WAVE ? nd : 0, WAVE);
        carry = (nd < WAVE) ? contCnt : 0u;
        qc += nd;
        c0 = c1;
    }
}

// A wave's share [ws, we) of an item's concatenated partner lists, two
// partners at a time: the records of 64 partners at a time are loaded one per
// lane (no dependent record load per partner), then the 64 lanes stream the
// elements of two consecutive partners together (U loads in flight per lane
// and partner) and probe them; a partner's count is a sum of ballot
// popcounts.  A partner cut by ws or we leaves its partial count in cutSet /
// cutCnt (slot 0: the one that begins before ws, which is also the one when
// the whole share lies inside a single partner; slot 1: the one that
// continues past we).  Partners of big rows are long (hundreds of elements on
// average), so whole-wave runs waste few lanes and need no
// position-to-partner search.  ADD: whole partners' counts are added to out
// (a partner counted in two universe halves) instead of stored.
template <int U, bool ADD, typename ProbeFn>
__device__ __forceinline__ void edge_runs(const uint4* __restrict__ part, uint32_t j0, uint32_t j1, uint32_t ws,
                                          uint32_t we, const uint32_t* __restrict__ elems, ProbeFn probeFn,
                                          uint32_t* __restrict__ out, uint32_t* cutSet, uint32_t* cutCnt) {
    const uint32_t lane = lane_id();
    const uint32_t lo = first_partner(part, j0, j1, ws);
    uint32_t qb = lo;
    uint32_t prevEnd = (qb == j0) ? 0u : part[qb - 1].z;
    while (qb < j1 && prevEnd < we) {
        const bool pv = qb + lane < j1;
        const uint4 R = pv ? part[qb + lane] : make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
        const uint32_t nrec = j1 - qb < uint32_t(WAVE) ? j1 - qb : uint32_t(WAVE);
        auto elemsOf = [&](uint32_t r) {
            const uint32_t offLo = uint32_t(__builtin_amdgcn_readlane(int(R.x), int(r)));
            const uint32_t offHi = uint32_t(__builtin_amdgcn_readlane(int(R.y), int(r)));
            return elems + ((uint64_t(offHi) << 32) | offLo);
        };
        // partner r's count for its elements [a, b) of its run [st, en)
        auto finish = [&](uint32_t r, uint32_t st, uint32_t en, uint32_t a, uint32_t b, uint32_t cnt) {
            if (lane == 0) {
                if (a == st && b == en) {
                    uint32_t* o = out + uint32_t(__builtin_amdgcn_readlane(int(R.w), int(r)));
                    if (ADD) atomicAdd(o, cnt); else *o = cnt;
                } else {
                    const int slot = a != st ? 0 : 1;
                    cutSet[slot] = qb + r;
                    cutCnt[slot] = cnt;
                }
            }
        };
        uint32_t start = prevEnd;
        for (uint32_t r = 0; r < nrec && start < we; r += 2) {
            // partner r covers [start, endA), partner r + 1 [endA, endB); both
            // have elements, and the wave's share begins inside partner r
            const uint32_t endA = uint32_t(__builtin_amdgcn_readlane(int(R.z), int(r)));
            const uint32_t aA = start > ws ? start : ws, bA = endA < we ? endA : we, nA = bA - aA;
            const uint32_t* srcA = elemsOf(r) + (aA - start);
            const bool hasB = r + 1 < nrec && endA < we;
            uint32_t endB = endA, bB = endA, nB = 0;
            const uint32_t* srcB = srcA;
            if (hasB) {
                endB = uint32_t(__builtin_amdgcn_readlane(int(R.z), int(r + 1)));
                bB = endB < we ? endB : we;
                nB = bB - endA;
                srcB = elemsOf(r + 1);
            }
            uint32_t cntA = 0, cntB = 0;
            const uint32_t n = nA > nB ? nA : nB;
            for (uint32_t i = 0; i < n; i += U * WAVE) {
                // straight-line loads: positions past a partner's end re-read
                // its last element (partner B's first when it has none)
                uint32_t xa[U], xb[U];
                bool va[U], vb[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t k = i + uint32_t(u) * WAVE + lane;
                    va[u] = k < nA;
                    vb[u] = k < nB;
                    xa[u] = srcA[va[u] ? k : nA - 1];
                    xb[u] = srcB[vb[u] ? k : (nB ? nB - 1 : 0u)];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    cntA += uint32_t(__popcll(__ballot(probeFn(xa[u]) & va[u])));
                    cntB += uint32_t(__popcll(__ballot(probeFn(xb[u]) & vb[u])));
                }
            }
            finish(r, start, endA, aA, bA, cntA);
            if (hasB) finish(r + 1, endA, endB, endA, bB, cntB);
            start = hasB ? endB : endA;
        }
        prevEnd = uint32_t(__builtin_amdgcn_readlane(int(R.z), int(nrec - 1)));
        qb += nrec;
    }
}

// Rows whose structure is shared by a workgroup (BITMAP, HASH, SEARCH):
// persistent workgroups pop items; the waves split the item's partner
// elements evenly and stream them partner by partner; partners cut by wave
// boundaries are completed through LDS.
// HALF: BITMAP items over one half of the universe (item mode bit 8: the
// upper half [32 * LDSW, 64 * LDSW)).  The row's elements in the half are
// [0, rowSplit[row]) or [rowSplit[row], n); the item's partner records
// already hold only the partners' elements in the half.  A 64 KiB bitmap
// lets two workgroups share a CU; each half adds its counts to out, which
// the caller zeroes first.
template <int NT, int LDSW, bool HALF = false>
__global__ __launch_bounds__(NT) void k_edges_wg(const Item* __restrict__ items, uint32_t nitems,
                                                 uint32_t* __restrict__ queue, const uint32_t* __restrict__ elems,
                                                 const uint64_t* __restrict__ offsets, const uint4* __restrict__ part,
                                                 uint32_t bmWords, const uint32_t* __restrict__ rowSplit,
                                                 uint32_t* __restrict__ out) {
    constexpr int NW = NT / WAVE;
    static_assert(NW <= 32, "partial slots");
    __shared__ RowSmem<LDSW> sm;
    const uint32_t tid = threadIdx.x, w = tid / WAVE;
    if (tid == 0) {
        sm.row = EMPTY;
        sm.mode = EMPTY;
    }
    __syncthreads();
    while (true) {
        if (tid == 0) sm.item = atomicAdd(queue, 1u);
        if (tid < 2 * NW) sm.pset[tid] = EMPTY;
        __syncthreads();
        const uint32_t it = sm.item;
        if (it >= nitems) break;
        const Item I = items[it];
        uint64_t rowOff = offsets[I.row], rowN = offsets[I.row + 1] - rowOff;
        uint32_t base = 0;
        if (HALF) {
            const uint32_t sp = rowSplit[I.row];
            if (I.mode >> 8) {
                rowOff += sp;
                rowN -= sp;
                base = uint32_t(LDSW) * 32u;
            } else {
                rowN = sp;
            }
        }
        if (I.row != sm.row || I.mode != sm.mode) {
            __syncthreads();
            build_row<NT, LDSW>(sm, elems + rowOff, rowN, HALF ? uint32_t(MODE_BITMAP) : I.mode, bmWords, base);
            if (tid == 0) {
                sm.row = I.row;
                sm.mode = I.mode;
            }
        }
        const uint32_t mode = HALF ? uint32_t(MODE_BITMAP) : I.mode, logTS = sm.logTS, hasMax = sm.hasMax;
        const uint32_t* S = elems + rowOff;
        const uint32_t T = part[I.j1 - 1].z;
        const uint32_t ws = uint32_t((uint64_t(T) * w) / NW), we = uint32_t((uint64_t(T) * (w + 1)) / NW);
        if (ws < we) {
            if (HALF || mode == MODE_BITMAP) {  // the common case, without the generic probe's branches
                edge_runs<4, HALF>(part, I.j0, I.j1, ws, we, elems,
                                   [&](uint32_t x) {
                                       const uint32_t y = x - base;
                                       return ((sm.tab[y >> 5] >> (y & 31)) & 1u) != 0u;
                                   },
                                   out, &sm.pset[2 * w], &sm.pcnt[2 * w]);
            } else if (!HALF) {
                edge_runs<4, false>(part, I.j0, I.j1, ws, we, elems,
                                    [&](uint32_t x) { return probe<LDSW>(sm, mode, x, S, rowN, logTS, hasMax) != 0u; },
                                    out, &sm.pset[2 * w], &sm.pcnt[2 * w]);
            }
        }
        __syncthreads();
        if (tid == 0) {
            for (int s = 0; s < 2 * NW; ++s) {
                const uint32_t j = sm.pset[s];
                if (j == EMPTY) continue;
                uint32_t tot = 0;
                for (int t = s; t < 2 * NW; ++t)
                    if (sm.pset[t] == j) {
                        tot += sm.pcnt[t];
                        if (t != s) sm.pset[t] = EMPTY;
                    }
                if (HALF) atomicAdd(&out[part[j].w], tot); else out[part[j].w] = tot;
            }
        }
        __syncthreads();
    }
}

// Small rows (at most WAVE_HASH_MAX elements): every wave pops its own items
// and keeps the row's open-addressing table in a private LDS region, so an
// item costs no workgroup barrier.
constexpr uint32_t WAVE_HASH_MAX = 511;
constexpr int WH_NT = 256, WH_SLOTS = 1024;

__global__ __launch_bounds__(WH_NT) void k_edges_wave(const Item* __restrict__ items, uint32_t nitems,
                                                      uint32_t* __restrict__ queue, const uint32_t* __restrict__ elems,
                                                      const uint64_t* __restrict__ offsets,
                                                      const uint4* __restrict__ part, uint32_t* __restrict__ out) {
    __shared__ uint32_t tabs[WH_NT / WAVE][WH_SLOTS];
    __shared__ uint64_t wmask[WH_NT / WAVE][VT + 1];  // store_chunk_masks, per wave
    __shared__ uint32_t wpre[WH_NT / WAVE][VT + 1];
    const uint32_t lane = lane_id();
    uint32_t* tab = tabs[threadIdx.x / WAVE];
    while (true) {
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(queue, 1u);
        it = __shfl(it, 0, WAVE);
        if (it >= nitems) break;
        const Item I = items[it];
        const uint64_t ro = offsets[I.row];
        const uint32_t rn = uint32_t(offsets[I.row + 1] - ro);
        uint32_t logTS = 6;
        while ((1u << logTS) < 4 * rn && (1u << logTS) < uint32_t(WH_SLOTS)) ++logTS;  // load <= 1/4 if it fits
        while ((1u << logTS) < 2 * rn) ++logTS;
        const uint32_t TS = 1u << logTS, mask = TS - 1;
        for (uint32_t i = lane; i < TS; i += WAVE) tab[i] = EMPTY;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool mx = false;
        for (uint32_t i = lane; i < rn; i += WAVE) {
            const uint32_t x = elems[ro + i];
            if (x == EMPTY) {
                mx = true;
                continue;
            }
            uint32_t sl = hslot(x, logTS);
            while (true) {
                const uint32_t prev = atomicCAS(&tab[sl], EMPTY, x);
                if (prev == EMPTY || prev == x) break;
                sl = (sl + 1) & mask;
            }
        }
        const uint32_t hasMax = __ballot(mx) ? 1u : 0u;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint32_t T = part[I.j1 - 1].z;
        if (T > 0)
            edge_batches(part, I.j0, I.j1, 0, T, elems,
                         [&](uint32_t x) -> uint32_t {
                             if (x == EMPTY) return hasMax;
                             uint32_t sl = hslot(x, logTS);
                             while (true) {
                                 const uint32_t k = tab[sl];
                                 if (k == x) return 1u;
                                 if (k == EMPTY) return 0u;
                                 sl = (sl + 1) & mask;
                             }
                         },
                         out, nullptr, nullptr, wmask[threadIdx.x / WAVE], wpre[threadIdx.x / WAVE]);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// out[idx[i]] = 0 for partners without elements
__global__ void k_zero_list(const uint32_t* __restrict__ idx, uint32_t n, uint32_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[idx[i]] = 0;
}

template <int NT>
__global__ __launch_bounds__(NT) void k_sum_counts(const uint32_t* __restrict__ c, uint64_t m,
                                                   uint64_t* __restrict__ partials, uint32_t* __restrict__ done,
                                                   uint64_t* __restrict__ out) {
    __shared__ uint64_t red[NT / WAVE];
    uint64_t acc = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * NT + threadIdx.x; i < m; i += uint64_t(gridDim.x) * NT) acc += c[i];
    grid_fold_u64<NT>(acc, partials, done, out, nullptr, red);
}

// Plain edge-wise counts, the baseline the binned dispatcher is measured
// against: one wave per edge (grid-stride), no grouping by vertex and no LDS
// tables.  The lanes take the elements of the shorter list and binary-search
// each in the longer list in global memory (lower bound over the whole list),
// the per-pair binary search of the reference's obs (obs.cu:330-345 searches
// each element of one set in the other); the wave sums the matches.
template <int NT>
__global__ __launch_bounds__(NT) void k_edges_plain(const uint2* __restrict__ edges, uint64_t m,
                                                    const uint32_t* __restrict__ elems,
                                                    const uint64_t* __restrict__ offs, uint32_t* __restrict__ counts,
                                                    uint64_t* __restrict__ partials, uint32_t* __restrict__ done,
                                                    uint64_t* __restrict__ total) {
    __shared__ uint64_t red[NT / WAVE];
    constexpr uint32_t WPB = NT / WAVE;
    const uint32_t lane = lane_id();
    uint64_t acc = 0;
    for (uint64_t e = uint64_t(blockIdx.x) * WPB + threadIdx.x / WAVE; e < m; e += uint64_t(gridDim.x) * WPB) {
        const uint2 uv = edges[e];
        const uint64_t oa = offs[uv.x], na = offs[uv.x + 1] - oa;
        const uint64_t ob = offs[uv.y], nb = offs[uv.y + 1] - ob;
        const bool aShort = na <= nb;
        const uint32_t* S = elems + (aShort ? oa : ob);
        const uint32_t* L = elems + (aShort ? ob : oa);
        const uint64_t ns = aShort ? na : nb, nl = aShort ? nb : na;
        uint64_t hits = 0;
        if (nl) {
            const uint64_t top = 1ull << (63 - __builtin_clzll(nl));  // largest power of two <= nl
            for (uint64_t i = lane; i < ns; i += WAVE) {
                const uint32_t x = S[i];
                uint64_t pos = 0;  // elements of L below x
                for (uint64_t step = top; step; step >>= 1)
                    if (pos + step <= nl && L[pos + step - 1] < x) pos += step;
                hits += uint64_t(pos < nl && L[pos] == x);
            }
        }
        const uint64_t c = wave_sum_u64(hits);
        if (lane == 0) {
            counts[e] = uint32_t(c);
            acc += c;
        }
    }
    grid_fold_u64<NT>(acc, partials, done, total, nullptr, red);
}

__global__ void k_reset_queues(uint32_t* q) {
    if (threadIdx.x < 5) q[threadIdx.x] = 0;
}

// lb[i] = the number of elements below key in list i (recs[i]: .x/.y element
// offset, .z length; lists sorted).
__global__ void k_lower_bounds(const uint4* __restrict__ recs, uint64_t n, const uint32_t* __restrict__ elems,
                               uint32_t key, uint32_t* __restrict__ lb) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint4 r = recs[i];
        const uint32_t* S = elems + ((uint64_t(r.y) << 32) | r.x);
        uint32_t lo = 0, hi = r.z;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (S[mid] < key) lo = mid + 1; else hi = mid;
        }
        lb[i] = lo;
    }
}

__global__ void k_fill_zero(uint32_t* __restrict__ out, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        out[i] = 0;
}

}  // namespace

// ------------------------------------------------------------------ host --
namespace si_impl {

struct BatchPlan {
    uint32_t k = 0, kne = 0;
    std::vector<uint32_t> perm, ssize;
    std::vector<uint64_t> soff;
    uint32_t* d_perm = nullptr;
    uint64_t* d_soff = nullptr;
    uint32_t* d_selem = nullptr;
    uint32_t* d_queue = nullptr;
    // items for the current tuning
    si_tuning tuned{};
    bool have_items = false;
    Item* d_bm_items = nullptr;
    Item* d_pr_items = nullptr;
    uint32_t* d_bm_starts = nullptr;  // per item, each wave's first partner set
    uint32_t* d_pr_starts = nullptr;
    uint4* d_merge = nullptr;
    uint32_t n_bm = 0, n_pr = 0, n_merge = 0;
    uint64_t items_cap_bm = 0, items_cap_pr = 0, merge_cap = 0, starts_cap_bm = 0, starts_cap_pr = 0;
    // captured launch sequence
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    void* g_out = nullptr;
    si_tuning g_tuning{};
    ~BatchPlan() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : {(void*)d_perm, (void*)d_soff, (void*)d_selem, (void*)d_queue, (void*)d_bm_items,
                        (void*)d_pr_items, (void*)d_bm_starts, (void*)d_pr_starts, (void*)d_merge})
            if (p) (void)hipFree(p);
    }
};

void free_batch_plan(BatchPlan* p) { delete p; }

static bool same_tuning(const si_tuning& a, const si_tuning& b) {
    return a.bitmap_min == b.bitmap_min && a.ratio_search == b.ratio_search && a.chunk_elems == b.chunk_elems &&
           a.chunk_elems_bitmap == b.chunk_elems_bitmap && a.probe_wg_per_cu == b.probe_wg_per_cu &&
           a.use_graph == b.use_graph;
}

static uint32_t bitmap_words(si_dataset* ds) { return uint32_t((uint64_t(ds->max_elem) >> 5) + 1); }

// Defaults measured on MI355X (the autotuner's picks): the all-pairs
// dispatcher on 16384 Zipf(1.2) sets, every row with a bitmap in 8 Mi
// partner elements per item (10.5 -> 9.4 ms against bitmap rows from 512
// elements and 2 Mi per item, since a bitmap probe costs four VALU
// instructions and a hash probe several times that), and the edge
// dispatcher on R-MAT 20 (bitmap rows from 512 elements, small hash-row
// items, two probe workgroups per CU: 10.6 -> 9.5 ms).
int batched_default_tuning(si_dataset*, si_tuning* t) {
    std::memset(t, 0, sizeof(*t));
    t->version = SI_TUNING_VERSION;
    t->bitmap_min = 16;
    t->ratio_search = 32;
    t->chunk_elems = 262144;
    t->chunk_elems_bitmap = 1u << 23;
    t->probe_wg_per_cu = 4;
    t->use_graph = 1;
    return SI_OK;
}

static void edge_default_tuning(si_dataset* ds, si_tuning* t) {
    batched_default_tuning(ds, t);
    t->bitmap_min = 512;
    t->chunk_elems = 16384;
    t->chunk_elems_bitmap = 1u << 21;
    t->probe_wg_per_cu = 2;
}

static int check_tuning(const si_tuning& t) {
    if (t.version != SI_TUNING_VERSION) return fail(SI_ERR_ARG, "tuning record has the wrong version");
    if (t.chunk_elems == 0 || t.chunk_elems_bitmap == 0 || t.probe_wg_per_cu == 0 || t.ratio_search == 0)
        return fail(SI_ERR_ARG, "tuning record has a zero field");
    return SI_OK;
}

static int build_plan(si_dataset* ds) {
    if (ds->plan) return SI_OK;
    si_context* ctx = ds->ctx;
    auto plan = std::make_unique<BatchPlan>();
    const uint32_t k = ds->k;
    plan->k = k;
    plan->perm.resize(k);
    std::iota(plan->perm.begin(), plan->perm.end(), 0u);
    std::stable_sort(plan->perm.begin(), plan->perm.end(),
                     [&](uint32_t a, uint32_t b) { return ds->sizes[a] > ds->sizes[b]; });
    plan->ssize.resize(k);
    plan->soff.assign(size_t(k) + 1, 0);
    for (uint32_t i = 0; i < k; ++i) {
        plan->ssize[i] = ds->sizes[plan->perm[i]];
        plan->soff[i + 1] = plan->soff[i] + plan->ssize[i];
    }
    plan->kne = 0;
    while (plan->kne < k && plan->ssize[plan->kne] > 0) ++plan->kne;
    SI_HIP(hipMalloc(&plan->d_perm, std::max<uint64_t>(k, 1) * 4));
    SI_HIP(hipMalloc(&plan->d_soff, (uint64_t(k) + 1) * 8));
    SI_HIP(hipMalloc(&plan->d_selem, (ds->total + SELEM_PAD) * 4));
    SI_HIP(hipMalloc(&plan->d_queue, 256));
    SI_HIP(hipMemcpyAsync(plan->d_perm, plan->perm.data(), uint64_t(k) * 4, hipMemcpyHostToDevice, ctx->stream));
    SI_HIP(hipMemcpyAsync(plan->d_soff, plan->soff.data(), (uint64_t(k) + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    SI_HIP(hipMemsetAsync(plan->d_selem + ds->total, 0, SELEM_PAD * 4, ctx->stream));
    SI_HIP(hipMemsetAsync(plan->d_queue, 0, 256, ctx->stream));
    if (k) {
        KernelScope ks(ctx, "batched_gather");
        hipLaunchKernelGGL(k_gather, dim3(std::min<uint32_t>(k, 65535)), dim3(256), 0, ctx->stream, ds->d_elems,
                           ds->d_offsets, plan->d_perm, plan->d_soff, k, plan->d_selem);
        SI_HIP(hipGetLastError());
    }
    SI_HIP(hipStreamSynchronize(ctx->stream));
    ds->plan = plan.release();
    return SI_OK;
}

// Builds the work items for `t` (host), uploads them.
static int build_items(si_dataset* ds, const si_tuning& t) {
    BatchPlan* P = ds->plan;
    if (P->have_items && same_tuning(P->tuned, t)) return SI_OK;
    si_context* ctx = ds->ctx;
    const uint32_t kne = P->kne;
    const bool bmFits = bitmap_words(ds) <= uint32_t(BM_WORDS);
    std::vector<uint4> merge;
    // A row's partners are the sets after it in size order, cut into items of
    // about `chunk` partner elements at set boundaries of one grid shared by
    // all rows, so that the items of different rows over the same grid chunk
    // read the same partner elements.  The items are ordered chunk-major: the
    // consecutive items that the resident workgroups pop together share
    // their partner elements in L2 instead of each streaming its own.
    auto grid = [&](uint64_t chunk) {
        std::vector<uint32_t> g{0};
        while (g.back() < kne) {
            const uint32_t j = g.back();
            const uint64_t target = P->soff[j] + chunk;
            uint32_t e = uint32_t(std::upper_bound(P->soff.begin() + j + 1, P->soff.begin() + kne + 1, target) -
                                  P->soff.begin()) - 1;
            if (e <= j) e = j + 1;  // one set larger than the chunk
            g.push_back(e);
        }
        return g;
    };
    const std::vector<uint32_t> gBm = grid(t.chunk_elems_bitmap), gPr = grid(t.chunk_elems);
    struct Tagged {
        uint32_t m;
        Item it;
    };
    std::vector<Tagged> bmT, prT;
    auto emit = [&](std::vector<Tagged>& v, const std::vector<uint32_t>& g, uint32_t row, uint32_t j0, uint32_t j1,
                    uint32_t mode) {
        uint32_t m = uint32_t(std::upper_bound(g.begin(), g.end(), j0) - g.begin()) - 1;  // g[m] <= j0 < g[m+1]
        for (uint32_t j = j0; j < j1; ++m) {
            const uint32_t e = std::min(g[m + 1], j1);
            v.push_back(Tagged{m, Item{row, j, e, mode}});
            j = e;
        }
    };
    for (uint32_t i = 0; i + 1 < kne; ++i) {
        const uint32_t si = P->ssize[i];
        if (bmFits && (si >= t.bitmap_min || si > HASH_MAX)) {
            emit(bmT, gBm, i, i + 1, kne, MODE_BITMAP);
        } else if (si <= HASH_MAX) {
            emit(prT, gPr, i, i + 1, kne, MODE_HASH);
        } else {
            // BIG row: partners with si >= ratio * |S_j| are searched, larger ones merged
            uint32_t j = i + 1;
            while (j < kne && uint64_t(P->ssize[j]) * t.ratio_search > si) {
                const uint32_t pa = std::min(P->perm[i], P->perm[j]), pb = std::max(P->perm[i], P->perm[j]);
                const uint64_t slot = (uint64_t(P->k) * (P->k - 1) / 2) - uint64_t(P->k - pa) * (P->k - pa - 1) / 2 + pb - pa - 1;
                merge.push_back(make_uint4(P->perm[i], P->perm[j], uint32_t(slot), uint32_t(slot >> 32)));
                ++j;
            }
            if (j < kne) emit(prT, gPr, i, j, kne, MODE_SEARCH);
        }
    }
    auto chunkMajor = [](std::vector<Tagged>& v) {
        std::stable_sort(v.begin(), v.end(), [](const Tagged& a, const Tagged& b) { return a.m < b.m; });
        std::vector<Item> out(v.size());
        for (size_t q = 0; q < v.size(); ++q) out[q] = v[q].it;
        return out;
    };
    std::vector<Item> bm = chunkMajor(bmT), pr = chunkMajor(prT);
    // each wave's first partner set: the largest j in [j0, j1) with
    // soff[j] <= the wave's first element (the waves split an item's partner
    // elements evenly, as the row kernel does)
    auto starts = [&](const std::vector<Item>& items, uint32_t nw) {
        std::vector<uint32_t> st(items.size() * nw);
        for (size_t q = 0; q < items.size(); ++q) {
            const Item& it = items[q];
            const uint64_t e0 = P->soff[it.j0], len = P->soff[it.j1] - e0;
            for (uint32_t w = 0; w < nw; ++w) {
                const uint64_t ws = e0 + (len * w) / nw;
                st[q * nw + w] = uint32_t(std::upper_bound(P->soff.begin() + it.j0 + 1, P->soff.begin() + it.j1, ws) -
                                          P->soff.begin()) - 1;
            }
        }
        return st;
    };
    std::vector<uint32_t> bmStarts = starts(bm, BM_NT / WAVE), prStarts = starts(pr, PR_NT / WAVE);
    auto upload = [&](auto& vec, auto*& dptr, uint64_t& cap) -> int {
        using T = typename std::decay<decltype(vec)>::type::value_type;
        if (vec.size() > cap || !dptr) {
            if (dptr) SI_HIP(hipFree(dptr));
            dptr = nullptr;
            cap = std::max<uint64_t>(vec.size(), 1);
            SI_HIP(hipMalloc(reinterpret_cast<void**>(&dptr), cap * sizeof(T)));
        }
        if (!vec.empty())
            SI_HIP(hipMemcpyAsync(dptr, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
        return SI_OK;
    };
    if (bm.size() > 0xFFFFFFFFull || pr.size() > 0xFFFFFFFFull || merge.size() > 0xFFFFFFFFull)
        return fail(SI_ERR_RANGE, "too many work items");
    SI_TRY(upload(bm, P->d_bm_items, P->items_cap_bm));
    SI_TRY(upload(pr, P->d_pr_items, P->items_cap_pr));
    SI_TRY(upload(bmStarts, P->d_bm_starts, P->starts_cap_bm));
    SI_TRY(upload(prStarts, P->d_pr_starts, P->starts_cap_pr));
    SI_TRY(upload(merge, P->d_merge, P->merge_cap));
    SI_HIP(hipStreamSynchronize(ctx->stream));
    P->n_bm = uint32_t(bm.size());
    P->n_pr = uint32_t(pr.size());
    P->n_merge = uint32_t(merge.size());
    P->tuned = t;
    P->have_items = true;
    if (P->gexec) {  // items changed: drop the captured sequence
        (void)hipGraphExecDestroy(P->gexec);
        (void)hipGraphDestroy(P->graph);
        P->gexec = nullptr;
        P->graph = nullptr;
    }
    return SI_OK;
}

// Enqueues the launch sequence of one batched count.
static int enqueue(si_dataset* ds, const si_tuning& t, uint32_t* out) {
    BatchPlan* P = ds->plan;
    si_context* ctx = ds->ctx;
    const uint32_t k = P->k;
    hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(64), 0, ctx->stream, P->d_queue);
    if (P->n_bm) {
        KernelScope ks(ctx, "batched_rows_bitmap");
        const uint32_t grid = uint32_t(std::min<uint64_t>(ctx->num_cus, P->n_bm));
        hipLaunchKernelGGL((k_rows<BM_NT, BM_WORDS>), dim3(grid), dim3(BM_NT), 0, ctx->stream, P->d_bm_items,
                           P->d_bm_starts, P->n_bm, P->d_queue + 0, P->d_selem, P->d_soff, P->d_perm, k,
                           bitmap_words(ds), out);
    }
    if (P->n_pr) {
        KernelScope ks(ctx, "batched_rows_probe");
        const uint32_t grid = uint32_t(std::min<uint64_t>(uint64_t(ctx->num_cus) * t.probe_wg_per_cu, P->n_pr));
        hipLaunchKernelGGL((k_rows<PR_NT, PR_WORDS>), dim3(grid), dim3(PR_NT), 0, ctx->stream, P->d_pr_items,
                           P->d_pr_starts, P->n_pr, P->d_queue + 1, P->d_selem, P->d_soff, P->d_perm, k,
                           bitmap_words(ds), out);
    }
    if (P->n_merge) {
        KernelScope ks(ctx, "batched_merge_pairs");
        const uint32_t grid = uint32_t(std::min<uint64_t>(uint64_t(ctx->num_cus) * 4, P->n_merge));
        hipLaunchKernelGGL((k_merge_pairs<256, 16>), dim3(grid), dim3(256), 0, ctx->stream, P->d_merge, P->n_merge,
                           ds->d_elems, ds->d_offsets, out);
    }
    if (P->kne < k && k > 1) {
        KernelScope ks(ctx, "batched_zero_pairs");
        hipLaunchKernelGGL(k_zero_pairs, dim3(std::min<uint32_t>(k, 4096)), dim3(256), 0, ctx->stream, P->d_perm, k,
                           P->kne, out);
    }
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int batched_count(si_dataset* ds, const si_tuning* tuning, uint32_t* d_out) {
    si_tuning t;
    if (tuning) {
        t = *tuning;
        SI_TRY(check_tuning(t));
    } else {
        batched_default_tuning(ds, &t);
    }
    if (ds->k < 2) return SI_OK;
    SI_TRY(build_plan(ds));
    SI_TRY(build_items(ds, t));
    BatchPlan* P = ds->plan;
    si_context* ctx = ds->ctx;
    // kernel timing records events between launches, which a graph cannot hold
    if (!t.use_graph || ctx->kernel_timing) return enqueue(ds, t, d_out);
    if (!P->gexec || P->g_out != d_out || !same_tuning(P->g_tuning, t)) {
        if (P->gexec) {
            (void)hipGraphExecDestroy(P->gexec);
            (void)hipGraphDestroy(P->graph);
            P->gexec = nullptr;
            P->graph = nullptr;
        }
        SI_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int st = enqueue(ds, t, d_out);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (st != SI_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return fail(SI_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        P->graph = g;
        SI_HIP(hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0));
        P->g_out = d_out;
        P->g_tuning = t;
    }
    SI_HIP(hipGraphLaunch(P->gexec, ctx->stream));
    return SI_OK;
}

// Candidate values of a knob around a seed value: seed / 4 .. seed * 4 in
// factors of two, clamped to [lo, hi].
static std::vector<uint32_t> around(uint32_t seed, uint32_t lo, uint32_t hi) {
    std::vector<uint32_t> v;
    for (int s = -2; s <= 2; ++s) {
        uint64_t x = s < 0 ? (uint64_t(seed) >> -s) : (uint64_t(seed) << s);
        x = std::max<uint64_t>(lo, std::min<uint64_t>(hi, x));
        if (std::find(v.begin(), v.end(), uint32_t(x)) == v.end()) v.push_back(uint32_t(x));
    }
    return v;
}

int batched_autotune(si_dataset* ds, double budget_ms, si_tuning* best) { return batched_autotune_seeded(ds, nullptr, budget_ms, best); }

int batched_autotune_seeded(si_dataset* ds, const si_tuning* seed, double budget_ms, si_tuning* best) {
    si_context* ctx = ds->ctx;
    si_tuning base;
    batched_default_tuning(ds, &base);
    if (seed) {
        SI_TRY(check_tuning(*seed));
        base = *seed;
    }
    base.use_graph = 0;
    *best = base;
    if (ds->k < 2) return SI_OK;
    const uint64_t slots = uint64_t(ds->k) * (ds->k - 1) / 2;
    void* out = nullptr;
    SI_HIP(hipMalloc(&out, slots * 4));
    hipEvent_t e0, e1;
    SI_HIP(hipEventCreate(&e0));
    SI_HIP(hipEventCreate(&e1));
    double spent = 0.0, bestMs = 1e300;
    auto measure = [&](const si_tuning& t, double& ms) -> int {
        SI_TRY(batched_count(ds, &t, static_cast<uint32_t*>(out)));  // builds items, warms caches
        SI_HIP(hipEventRecord(e0, ctx->stream));
        SI_TRY(batched_count(ds, &t, static_cast<uint32_t*>(out)));
        SI_HIP(hipEventRecord(e1, ctx->stream));
        SI_HIP(hipEventSynchronize(e1));
        float f = 0.f;
        SI_HIP(hipEventElapsedTime(&f, e0, e1));
        ms = f;
        spent += 2.0 * f;
        return SI_OK;
    };
    int st = SI_OK;
    double ms = 0;
    st = measure(base, ms);
    if (st == SI_OK) bestMs = ms;
    // coordinate descent over the dispatcher's knobs: fixed candidate lists,
    // or values around the seed's when the descent starts from a seed
    std::vector<std::pair<uint32_t si_tuning::*, std::vector<uint32_t>>> knobs = {
        {&si_tuning::bitmap_min, {16, 64, 256, 512, 1024, 2048, 4096}},
        {&si_tuning::chunk_elems, {16384, 32768, 65536, 131072, 262144}},
        {&si_tuning::chunk_elems_bitmap, {1u << 20, 1u << 21, 1u << 22, 1u << 23, 1u << 24}},
        {&si_tuning::probe_wg_per_cu, {2, 4, 6, 8}},
        {&si_tuning::ratio_search, {8, 16, 32, 64}},
    };
    if (seed) {
        knobs[0].second = around(base.bitmap_min, 16, 1u << 16);
        knobs[1].second = around(base.chunk_elems, 4096, 1u << 22);
        knobs[2].second = around(base.chunk_elems_bitmap, 1u << 16, 1u << 24);
        knobs[3].second = {std::max(1u, base.probe_wg_per_cu / 2), base.probe_wg_per_cu, base.probe_wg_per_cu + 2,
                           base.probe_wg_per_cu * 2};
        knobs[4].second = around(base.ratio_search, 2, 1024);
    }
    for (int pass = 0; pass < 2 && st == SI_OK; ++pass) {
        for (auto& kv : knobs) {
            for (uint32_t v : kv.second) {
                if (st != SI_OK || spent > budget_ms) break;
                si_tuning t = *best;
                if (t.*(kv.first) == v) continue;
                t.*(kv.first) = v;
                st = measure(t, ms);
                if (st == SI_OK && ms < bestMs) {
                    bestMs = ms;
                    *best = t;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    best->use_graph = 1;
    return st;
}

// ------------------------------------------------------------- edge lists --
}  // namespace si_impl

struct si_edge_plan {
    si_dataset* ds = nullptr;
    uint64_t m = 0;
    si_tuning t{};
    Item* d_wg = nullptr;    // BITMAP / HASH / SEARCH rows (workgroup items)
    Item* d_wave = nullptr;  // small HASH rows (wave items)
    Item* d_half = nullptr;  // BITMAP rows by universe half (mode bit 8: upper half)
    uint4* d_part = nullptr;
    uint4* d_hpart = nullptr;       // partner records of d_half items, elements in the item's half
    uint32_t* d_rowsplit = nullptr;  // per row: its elements below the universe midpoint
    uint4* d_merge = nullptr;
    uint32_t* d_zero = nullptr;
    uint32_t* d_queue = nullptr;
    uint64_t* d_partials = nullptr;
    uint2* d_edges = nullptr;  // (u, v) per edge, for the plain baseline
    uint32_t n_bm = 0, n_wg = 0, n_wave = 0, n_merge = 0, n_zero = 0, n_half = 0;
    bool bm64 = false;  // whole-universe bitmaps fit 64 KiB: two bitmap workgroups per CU
    ~si_edge_plan() {
        for (void* p : {(void*)d_wg, (void*)d_wave, (void*)d_half, (void*)d_part, (void*)d_hpart,
                        (void*)d_rowsplit, (void*)d_merge, (void*)d_zero, (void*)d_queue, (void*)d_partials,
                        (void*)d_edges})
            if (p) (void)hipFree(p);
    }
};

namespace si_impl {

constexpr int SUM_GRID = 1024;

int edge_plan_create(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m, const si_tuning* tuning,
                     si_edge_plan** out) {
    si_tuning t;
    if (tuning) {
        t = *tuning;
        SI_TRY(check_tuning(t));
    } else {
        edge_default_tuning(ds, &t);
    }
    const uint32_t k = ds->k;
    if (m > 0xFFFFFFFFull) return fail(SI_ERR_RANGE, "at most 2^32-1 edges per plan");
    for (uint64_t e = 0; e < m; ++e)
        if (u[e] >= k || v[e] >= k) return fail(SI_ERR_ARG, "edge endpoint out of range (edge " + std::to_string(e) + ")");
    auto plan = std::make_unique<si_edge_plan>();
    plan->ds = ds;
    plan->m = m;
    plan->t = t;
    const auto& sz = ds->sizes;
    const auto& of = ds->offsets;
    // row = endpoint with the larger list (ties: the smaller id); counting sort by row
    std::vector<uint32_t> rowOf(m), cnt(size_t(k) + 1, 0);
    for (uint64_t e = 0; e < m; ++e) {
        const uint32_t a = u[e], b = v[e];
        const uint32_t r = (sz[a] > sz[b] || (sz[a] == sz[b] && a <= b)) ? a : b;
        rowOf[e] = r;
        cnt[r + 1]++;
    }
    for (uint32_t i = 0; i < k; ++i) cnt[i + 1] += cnt[i];
    std::vector<uint2> plist(m);  // (partner set, edge index) grouped by row
    {
        std::vector<uint32_t> pos(cnt.begin(), cnt.end() - 1);
        for (uint64_t e = 0; e < m; ++e) {
            const uint32_t r = rowOf[e];
            plist[pos[r]++] = make_uint2(u[e] == r ? v[e] : u[e], uint32_t(e));
        }
    }
    const bool bmFits = bitmap_words(ds) <= uint32_t(BM_WORDS);
    // universe halves: bitmap rows split at 32 * HB_WORDS, each half an item of its own
    const bool halves = bmFits && t.edge_bitmap_halves && bitmap_words(ds) > uint32_t(HB_WORDS);
    plan->bm64 = t.edge_bitmap_halves && bitmap_words(ds) <= uint32_t(HB_WORDS);
    std::vector<Item> bmItems, wgItems, waveItems, halfItems;
    std::vector<uint4> part, merge, hraw, hpart;
    std::vector<uint32_t> zero;
    std::vector<uint3> hrows;  // (row, first, end) of its records in hraw
    part.reserve(m);
    // partners [p0, p1) of `pv` (already appended) cut into items of about
    // `chunk` partner elements; pv[q].z becomes the inclusive end within its item
    auto emit = [&](std::vector<uint4>& pv, std::vector<Item>& vec, uint32_t row, uint32_t p0, uint32_t p1,
                    uint32_t mode, uint64_t chunk) {
        uint32_t p = p0;
        while (p < p1) {
            uint64_t acc = 0;
            uint32_t e = p;
            while (e < p1 && (e == p || acc + pv[e].z <= chunk)) acc += pv[e++].z;
            uint32_t run = 0;
            for (uint32_t q = p; q < e; ++q) {
                run += pv[q].z;
                pv[q].z = run;
            }
            vec.push_back(Item{row, p, e, mode});
            p = e;
        }
    };
    auto add = [&](const uint2& pe) {
        const uint32_t n = sz[pe.x];
        if (n == 0) {
            zero.push_back(pe.y);
            return;
        }
        const uint64_t o = of[pe.x];
        part.push_back(make_uint4(uint32_t(o), uint32_t(o >> 32), n, pe.y));  // .z: size until emit
    };
    for (uint32_t r = 0; r < k; ++r) {
        if (cnt[r] == cnt[r + 1]) continue;
        const uint32_t s = sz[r];
        const uint32_t start = uint32_t(part.size());
        const bool useBm = bmFits && (s >= t.bitmap_min || s > HASH_MAX);
        if (s == 0) {
            for (uint32_t q = cnt[r]; q < cnt[r + 1]; ++q) zero.push_back(plist[q].y);
            continue;
        }
        if (useBm && halves) {
            const uint32_t h0 = uint32_t(hraw.size());
            for (uint32_t q = cnt[r]; q < cnt[r + 1]; ++q) {
                const uint2 pe = plist[q];
                const uint32_t n = sz[pe.x];
                if (n == 0) {
                    zero.push_back(pe.y);
                    continue;
                }
                const uint64_t o = of[pe.x];
                hraw.push_back(make_uint4(uint32_t(o), uint32_t(o >> 32), n, pe.y));
            }
            if (uint32_t(hraw.size()) > h0) hrows.push_back(make_uint3(r, h0, uint32_t(hraw.size())));
        } else if (useBm || s <= HASH_MAX) {
            for (uint32_t q = cnt[r]; q < cnt[r + 1]; ++q) add(plist[q]);
            const uint32_t end = uint32_t(part.size());
            if (end == start) continue;
            if (useBm)
                emit(part, bmItems, r, start, end, MODE_BITMAP, t.chunk_elems_bitmap);
            else if (s <= WAVE_HASH_MAX)
                emit(part, waveItems, r, start, end, MODE_HASH, t.chunk_elems);
            else
                emit(part, wgItems, r, start, end, MODE_HASH, t.chunk_elems);
        } else {
            for (uint32_t q = cnt[r]; q < cnt[r + 1]; ++q) {
                const uint2 pe = plist[q];
                if (sz[pe.x] == 0)
                    zero.push_back(pe.y);
                else if (uint64_t(sz[pe.x]) * t.ratio_search > s)
                    merge.push_back(make_uint4(r, pe.x, pe.y, 0));
                else
                    add(pe);
            }
            const uint32_t end = uint32_t(part.size());
            if (end > start) emit(part, wgItems, r, start, end, MODE_SEARCH, t.chunk_elems);
        }
    }
    si_context* ctx = ds->ctx;
    std::vector<uint32_t> rowSplit;
    if (!hrows.empty()) {
        // each list's elements below the midpoint, by a device binary search
        // (partner lists, then the rows)
        std::vector<uint4> recs(hraw);
        for (const uint3& h : hrows) {
            const uint64_t o = of[h.x];
            recs.push_back(make_uint4(uint32_t(o), uint32_t(o >> 32), sz[h.x], 0u));
        }
        const uint64_t nrec = recs.size();
        uint4* d_recs = nullptr;
        uint32_t* d_lb = nullptr;
        std::vector<uint32_t> lb(nrec);
        int st = SI_OK;
        if (hipMalloc(reinterpret_cast<void**>(&d_recs), nrec * sizeof(uint4)) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&d_lb), nrec * 4) != hipSuccess)
            st = fail(SI_ERR_NOMEM, "edge plan: device memory for the universe-half split");
        if (st == SI_OK &&
            hipMemcpyAsync(d_recs, recs.data(), nrec * sizeof(uint4), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            st = fail(SI_ERR_HIP, "edge plan: upload of the split records");
        if (st == SI_OK) {
            const uint32_t grid = uint32_t(std::min<uint64_t>((nrec + 255) / 256, 4096));
            hipLaunchKernelGGL(k_lower_bounds, dim3(grid), dim3(256), 0, ctx->stream, d_recs, nrec, ds->d_elems,
                               uint32_t(HB_WORDS) * 32u, d_lb);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(lb.data(), d_lb, nrec * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                st = fail(SI_ERR_HIP, "edge plan: universe-half split");
        }
        if (d_recs) (void)hipFree(d_recs);
        if (d_lb) (void)hipFree(d_lb);
        SI_TRY(st);
        rowSplit.assign(k, 0u);
        hpart.reserve(2 * hraw.size());
        for (size_t i = 0; i < hrows.size(); ++i) {
            const uint3 h = hrows[i];
            rowSplit[h.x] = lb[hraw.size() + i];
            for (uint32_t half = 0; half < 2; ++half) {
                const uint32_t start = uint32_t(hpart.size());
                for (uint32_t q = h.y; q < h.z; ++q) {
                    const uint4 R = hraw[q];
                    const uint32_t sp = lb[q], nh = half ? R.z - sp : sp;
                    if (nh == 0) continue;  // no elements in this half: adds 0 to a zeroed count
                    const uint64_t o = ((uint64_t(R.y) << 32) | R.x) + (half ? sp : 0u);
                    hpart.push_back(make_uint4(uint32_t(o), uint32_t(o >> 32), nh, R.w));
                }
                const uint32_t end = uint32_t(hpart.size());
                if (end > start)
                    emit(hpart, halfItems, h.x, start, end, MODE_BITMAP | (half << 8), t.chunk_elems_bitmap);
            }
        }
    }
    auto up = [&](auto& vec, auto*& dptr) -> int {
        using T = typename std::decay<decltype(vec)>::type::value_type;
        SI_HIP(hipMalloc(reinterpret_cast<void**>(&dptr), std::max<size_t>(vec.size(), 1) * sizeof(T)));
        if (!vec.empty())
            SI_HIP(hipMemcpyAsync(dptr, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
        return SI_OK;
    };
    // workgroup items: BITMAP first (launched with the 1024-thread kernel), then the rest
    plan->n_bm = uint32_t(bmItems.size());
    bmItems.insert(bmItems.end(), wgItems.begin(), wgItems.end());
    SI_TRY(up(bmItems, plan->d_wg));
    SI_TRY(up(waveItems, plan->d_wave));
    SI_TRY(up(part, plan->d_part));
    SI_TRY(up(halfItems, plan->d_half));
    SI_TRY(up(hpart, plan->d_hpart));
    SI_TRY(up(rowSplit, plan->d_rowsplit));
    SI_TRY(up(merge, plan->d_merge));
    SI_TRY(up(zero, plan->d_zero));
    SI_HIP(hipMalloc(&plan->d_queue, 256));
    SI_HIP(hipMemsetAsync(plan->d_queue, 0, 256, ctx->stream));
    SI_HIP(hipMalloc(&plan->d_partials, SUM_GRID * 8));
    {
        std::vector<uint2> edges(m);
        for (uint64_t e = 0; e < m; ++e) edges[e] = make_uint2(u[e], v[e]);
        SI_TRY(up(edges, plan->d_edges));
        SI_HIP(hipStreamSynchronize(ctx->stream));  // before `edges` goes out of scope
    }
    plan->n_wg = uint32_t(wgItems.size());
    plan->n_wave = uint32_t(waveItems.size());
    plan->n_merge = uint32_t(merge.size());
    plan->n_zero = uint32_t(zero.size());
    plan->n_half = uint32_t(halfItems.size());
    ds->ctx->edge_plans.insert(plan.get());
    *out = plan.release();
    return SI_OK;
}

int edge_count(si_edge_plan* P, uint32_t* counts, uint64_t* total) {
    si_dataset* ds = P->ds;
    si_context* ctx = ds->ctx;
    hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(64), 0, ctx->stream, P->d_queue);
    if (P->n_half) {
        // universe halves add their counts: every count starts at zero (this
        // also covers the empty-list edges)
        const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((P->m + 255) / 256, 4096)));
        hipLaunchKernelGGL(k_fill_zero, dim3(grid), dim3(256), 0, ctx->stream, counts, P->m);
    } else if (P->n_zero) {
        const uint32_t grid = uint32_t(std::min<uint64_t>((P->n_zero + 255) / 256, 1024));
        hipLaunchKernelGGL(k_zero_list, dim3(grid), dim3(256), 0, ctx->stream, P->d_zero, P->n_zero, counts);
    }
    // the small-row kernel runs on the side stream beside the big-row ones:
    // a bitmap workgroup holds 128 KiB of LDS and 16 waves of a CU, which
    // leaves room for wave-kernel workgroups next to it
    const bool forked = P->n_wave && (P->n_bm || P->n_wg);
    if (P->n_wave) {
        hipStream_t s = ctx->stream;
        if (forked) {
            SI_HIP(hipEventRecord(ctx->fork, ctx->stream));
            SI_HIP(hipStreamWaitEvent(ctx->side, ctx->fork, 0));
            s = ctx->side;
        }
        KernelScope ks(ctx, "edges_rows_wave", s);
        const uint32_t grid = uint32_t(std::min<uint64_t>(uint64_t(ctx->num_cus) * 8, (P->n_wave + 3) / 4));
        hipLaunchKernelGGL(k_edges_wave, dim3(grid), dim3(WH_NT), 0, s, P->d_wave, P->n_wave, P->d_queue + 3,
                           ds->d_elems, ds->d_offsets, P->d_part, counts);
    }
    if (P->n_half) {
        KernelScope ks(ctx, "edges_rows_bitmap_half");
        const uint32_t grid = uint32_t(std::min<uint64_t>(uint64_t(ctx->num_cus) * 2, P->n_half));
        hipLaunchKernelGGL((k_edges_wg<BM_NT, HB_WORDS, true>), dim3(grid), dim3(BM_NT), 0, ctx->stream, P->d_half,
                           P->n_half, P->d_queue + 4, ds->d_elems, ds->d_offsets, P->d_hpart, uint32_t(HB_WORDS),
                           P->d_rowsplit, counts);
    }
    if (P->n_bm && P->bm64) {
        KernelScope ks(ctx, "edges_rows_bitmap");
        const uint32_t grid = uint32_t(std::min<uint64_t>(uint64_t(ctx->num_cus) * 2, P->n_bm));
        hipLaunchKernelGGL((k_edges_wg<BM_NT, HB_WORDS>), dim3(grid), dim3(BM_NT), 0, ctx->stream, P->d_wg, P->n_bm,
                           P->d_queue + 0, ds->d_elems, ds->d_offsets, P->d_part, bitmap_words(ds), nullptr, counts);
    } else if (P->n_bm) {
        KernelScope ks(ctx, "edges_rows_bitmap");
        const uint32_t grid = uint32_t(std::min<uint64_t>(ctx->num_cus, P->n_bm));
        hipLaunchKernelGGL((k_edges_wg<BM_NT, BM_WORDS>), dim3(grid), dim3(BM_NT), 0, ctx->stream, P->d_wg, P->n_bm,
                           P->d_queue + 0, ds->d_elems, ds->d_offsets, P->d_part, bitmap_words(ds), nullptr, counts);
    }
    if (P->n_wg) {
        KernelScope ks(ctx, "edges_rows_probe");
        const uint32_t grid = uint32_t(std::min<uint64_t>(uint64_t(ctx->num_cus) * P->t.probe_wg_per_cu, P->n_wg));
        hipLaunchKernelGGL((k_edges_wg<PR_NT, PR_WORDS>), dim3(grid), dim3(PR_NT), 0, ctx->stream, P->d_wg + P->n_bm,
                           P->n_wg, P->d_queue + 1, ds->d_elems, ds->d_offsets, P->d_part, bitmap_words(ds), nullptr,
                           counts);
    }
    if (forked) {
        SI_HIP(hipEventRecord(ctx->join, ctx->side));
        SI_HIP(hipStreamWaitEvent(ctx->stream, ctx->join, 0));
    }
    if (P->n_merge) {
        KernelScope ks(ctx, "edges_merge_pairs");
        const uint32_t grid = uint32_t(std::min<uint64_t>(uint64_t(ctx->num_cus) * 4, P->n_merge));
        hipLaunchKernelGGL((k_merge_pairs<256, 16>), dim3(grid), dim3(256), 0, ctx->stream, P->d_merge, P->n_merge,
                           ds->d_elems, ds->d_offsets, counts);
    }
    {
        KernelScope ks(ctx, "edges_sum");
        const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(SUM_GRID, (P->m + 255) / 256)));
        hipLaunchKernelGGL((k_sum_counts<256>), dim3(grid), dim3(256), 0, ctx->stream, counts, P->m, P->d_partials,
                           P->d_queue + 2, total);
    }
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int edge_count_plain(si_edge_plan* P, uint32_t* counts, uint64_t* total) {
    si_dataset* ds = P->ds;
    si_context* ctx = ds->ctx;
    KernelScope ks(ctx, "edges_plain");
    // one partial per workgroup: at most SUM_GRID workgroups
    const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(SUM_GRID, (P->m + 3) / 4)));
    hipLaunchKernelGGL((k_edges_plain<256>), dim3(grid), dim3(256), 0, ctx->stream, P->d_edges, P->m, ds->d_elems,
                       ds->d_offsets, counts, P->d_partials, P->d_queue + 2, total);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int edge_autotune(si_dataset* ds, const uint32_t* u, const uint32_t* v, uint64_t m, double budget_ms,
                  si_tuning* best) {
    si_context* ctx = ds->ctx;
    edge_default_tuning(ds, best);
    best->use_graph = 0;
    if (m == 0) return SI_OK;
    const auto t0 = std::chrono::steady_clock::now();
    auto wallMs = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    void* counts = nullptr;
    uint64_t* total = nullptr;
    SI_HIP(hipMalloc(&counts, m * 4));
    SI_HIP(hipMalloc(reinterpret_cast<void**>(&total), 8));
    hipEvent_t e0, e1;
    SI_HIP(hipEventCreate(&e0));
    SI_HIP(hipEventCreate(&e1));
    // plan with t, one warm run, the faster of two timed runs (single runs
    // differ by a few percent from run to run)
    auto measure = [&](const si_tuning& t, double& ms) -> int {
        si_edge_plan* plan = nullptr;
        SI_TRY(edge_plan_create(ds, u, v, m, &t, &plan));
        int st = edge_count(plan, static_cast<uint32_t*>(counts), total);
        ms = 1e300;
        for (int rep = 0; rep < 2 && st == SI_OK; ++rep) {
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) st = SI_ERR_HIP;
            if (st == SI_OK) st = edge_count(plan, static_cast<uint32_t*>(counts), total);
            if (st == SI_OK && hipEventRecord(e1, ctx->stream) != hipSuccess) st = SI_ERR_HIP;
            float f = 0.f;
            if (st == SI_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&f, e0, e1) != hipSuccess))
                st = fail(SI_ERR_HIP, "edge autotune timing failed");
            if (st == SI_OK && f < ms) ms = f;
        }
        edge_plan_free(plan);
        return st;
    };
    double bestMs = 0.0;
    int st = measure(*best, bestMs);
    // one pass of coordinate descent over the knobs that shape the edge plan
    const std::vector<std::pair<uint32_t si_tuning::*, std::vector<uint32_t>>> knobs = {
        {&si_tuning::bitmap_min, {64, 256, 1024, 2048}},
        {&si_tuning::chunk_elems_bitmap, {1u << 19, 1u << 23}},
        {&si_tuning::chunk_elems, {8192, 32768}},
        {&si_tuning::probe_wg_per_cu, {4, 8}},
        {&si_tuning::ratio_search, {16, 64}},
        {&si_tuning::edge_bitmap_halves, {0, 1}},
    };
    for (auto& kv : knobs) {
        for (uint32_t val : kv.second) {
            if (st != SI_OK || wallMs() > budget_ms) break;
            si_tuning t = *best;
            if (t.*(kv.first) == val) continue;
            t.*(kv.first) = val;
            double ms = 0.0;
            st = measure(t, ms);
            if (st == SI_OK && ms < bestMs) {
                bestMs = ms;
                *best = t;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(counts);
    (void)hipFree(total);
    return st;
}

void edge_plan_free(si_edge_plan* p) {
    p->ds->ctx->edge_plans.erase(p);
    (void)hipStreamSynchronize(p->ds->ctx->stream);
    delete p;
}

uint64_t plan_edges(si_edge_plan* p) { return p->m; }

si_dataset* plan_dataset(si_edge_plan* p) { return p->ds; }

}  // namespace si_impl
