// Device-side building blocks shared by the gfx950 kernels: wave64 reductions
// and scans, merge-path co-rank searches, and pair-index decoding.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace si_dev {

constexpr int WAVE = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// a wave-uniform value into a scalar register (first active lane's copy)
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return uint32_t(__builtin_amdgcn_readfirstlane(int(v))); }
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    return (uint64_t(uniform_u32(uint32_t(v >> 32))) << 32) | uniform_u32(uint32_t(v));
}

// Wave issue priority (s_setprio takes an immediate).  Kernels that stream
// one long range per workgroup, several workgroups per CU, rotate it per
// tile or bucket (priority (step + seed) mod 4): a CU otherwise issues from
// its oldest waves first, so the first of its workgroups runs ahead and the
// launch ends with the last ones (SISI merge at 2^28 x 2^28: materialise
// 0.553 -> 0.534 ms, count 0.441 -> 0.420 ms).
constexpr uint32_t kNoPrio = 0xffffffffu;
__device__ __forceinline__ void set_wave_prio(uint32_t p) {
    switch (p) {
        case 0: __builtin_amdgcn_s_setprio(0); break;
        case 1: __builtin_amdgcn_s_setprio(1); break;
        case 2: __builtin_amdgcn_s_setprio(2); break;
        default: __builtin_amdgcn_s_setprio(3); break;
    }
}

// --------------------------------------------------------------- reductions
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Inclusive prefix sum across the wave.
__device__ __forceinline__ uint32_t wave_inclusive_scan_u32(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        uint32_t t = __shfl_up(v, off, WAVE);
        if (lane >= uint32_t(off)) v += t;
    }
    return v;
}

// Exclusive prefix sum over the wave of values below 2^BITS, one bit plane at
// a time (ballot + mbcnt: no cross-lane data movement): lane l receives the
// sum over lanes below l, `total` the sum over the wave.
template <int BITS>
__device__ __forceinline__ uint32_t wave_excl_scan_bits(uint32_t v, uint32_t& total) {
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const uint64_t m = __ballot((v >> b) & 1u);
        pre += __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u)) << b;
        tot += uint32_t(__popcll(m)) << b;
    }
    total = tot;
    return pre;
}

// Sum over a workgroup of NT threads; `scratch` holds NT/64 uint64 slots.
// Every thread receives the total.  Contains two barriers.
template <int NT>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* scratch) {
    v = wave_sum_u64(v);
    const int w = threadIdx.x / WAVE;
    if (NT == WAVE) return v;
    if (lane_id() == 0) scratch[w] = v;
    __syncthreads();
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < NT / WAVE; ++i) t += scratch[i];
    __syncthreads();
    return t;
}

// Grid-wide sum of one uint64 per thread without atomics on the result:
// every workgroup stores its partial, takes a ticket, and the last one folds
// the partials into *out64 (and *out32 when given) and resets the ticket to
// zero, so the counter needs no clearing between launches.
template <int NT>
__device__ __forceinline__ void grid_fold_u64(uint64_t v, uint64_t* __restrict__ partials, uint32_t* __restrict__ done,
                                              uint64_t* __restrict__ out64, uint32_t* __restrict__ out32,
                                              uint64_t* scratch) {
    __shared__ uint32_t isLast;
    const uint64_t total = block_sum_u64<NT>(v, scratch);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = total;
        __threadfence();
        const uint32_t ticket = atomicAdd(done, 1u);
        isLast = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (!isLast) return;
    __threadfence();
    uint64_t acc = 0;
    for (uint32_t i = threadIdx.x; i < gridDim.x; i += NT)
        acc += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    acc = block_sum_u64<NT>(acc, scratch);
    if (threadIdx.x == 0) {
        if (out64) *out64 = acc;
        if (out32) *out32 = uint32_t(acc);
        *done = 0;
    }
}

// ------------------------------------------------------------- co-rank search
// Merge order used everywhere: A[x] precedes B[y] iff A[x] <= B[y] (ties put
// the A element first).  corank(d) = number of A elements among the first d
// merged elements = the smallest i in [max(0,d-nB), min(d,nA)] with
// A[i] > B[d-1-i] (or the upper bound if there is none).
//
// With this order the matching partner of A[x] is always the B element that
// is next in line when A[x] is merged, so an A element taken inside any
// partition of the merge path only ever needs the B "head" that follows the
// partition's B range.  No tie can be split between partitions.

// Half-wave (32 lanes) cooperative co-rank search in global memory: 32-ary
// search, ceil(log32(n)) dependent rounds.  `half` selects lanes 0-31 or
// 32-63; both halves may search different diagonals concurrently.  All lanes
// of the half must call it with the same arguments.
__device__ __forceinline__ uint64_t corank_halfwave(const uint32_t* __restrict__ A, uint64_t nA,
                                                    const uint32_t* __restrict__ B, uint64_t nB, uint64_t d) {
    const uint32_t lane = lane_id() & 31u;
    const uint32_t half = lane_id() >> 5;
    uint64_t lo = d > nB ? d - nB : 0;
    uint64_t hi = d < nA ? d : nA;
    while (lo < hi) {
        const uint64_t len = hi - lo;
        const uint64_t step = (len + 31) >> 5;
        const uint64_t p = lo + uint64_t(lane) * step;
        bool pred = false;
        if (p < hi) pred = A[p] > B[d - 1 - p];
        const uint64_t m64 = __ballot(pred);
        const uint32_t m = uint32_t(m64 >> (32 * half));
        if (m) {
            const uint32_t f = __builtin_ctz(m);
            const uint64_t pf = lo + uint64_t(f) * step;
            const uint64_t newlo = f ? lo + uint64_t(f - 1) * step + 1 : lo;
            lo = newlo;
            hi = pf;
        } else {
            const uint64_t last = (hi - 1 - lo) / step;  // last probed lane (<= 31)
            lo = lo + last * step + 1;
        }
    }
    return lo;
}

// corank_halfwave restricted to [lo, hi]: the caller knows the co-rank lies
// there (e.g. the end of a segment whose start co-rank is known: the end of
// a segment of L merged elements starting at (ia, ib) lies in
// [max(ia, d - nB), min(nA, ia + L)]), so the search takes ceil(log32(hi-lo))
// rounds.  Same calling rules as corank_halfwave.
__device__ __forceinline__ uint64_t corank_halfwave_range(const uint32_t* __restrict__ A, const uint32_t* __restrict__ B,
                                                          uint64_t d, uint64_t lo, uint64_t hi) {
    const uint32_t lane = lane_id() & 31u;
    const uint32_t half = lane_id() >> 5;
    while (lo < hi) {
        const uint64_t step = (hi - lo + 31) >> 5;
        const uint64_t p = lo + uint64_t(lane) * step;
        const bool pred = p < hi && A[p] > B[d - 1 - p];
        const uint32_t m = uint32_t(__ballot(pred) >> (32 * half));
        if (m) {
            const uint32_t f = __builtin_ctz(m);
            const uint64_t pf = lo + uint64_t(f) * step;
            lo = f ? lo + uint64_t(f - 1) * step + 1 : lo;
            hi = pf;
        } else {
            lo = lo + ((hi - 1 - lo) / step) * step + 1;
        }
    }
    return lo;
}

// Single-thread binary co-rank search (used for short ranges).
__device__ __forceinline__ uint64_t corank_serial(const uint32_t* __restrict__ A, uint64_t nA,
                                                  const uint32_t* __restrict__ B, uint64_t nB, uint64_t d) {
    uint64_t lo = d > nB ? d - nB : 0;
    uint64_t hi = d < nA ? d : nA;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (A[mid] <= B[d - 1 - mid])
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// ------------------------------------------------- sample-cache search (obs)
// Searches x in the sorted set S (size nS) whose samples
// cache[t] = S[t*nS/NS], t < NS, are in LDS; returns 1 if present.  Phase 1
// brackets x between two samples in LDS, phase 2 binary-searches S inside
// the bracket in global memory (obs.cu:312-345), with 64-bit positions.
template <int NS>
__device__ __forceinline__ uint32_t sample_search(const uint32_t* __restrict__ S, uint64_t nS, const uint32_t* cache,
                                                  uint32_t x) {
    uint32_t lo = 0, hi = NS;  // phase 1: lo = number of samples <= x
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cache[mid] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo == 0) return 0;  // x < S[0]
    if (cache[lo - 1] == x) return 1;
    // phase 2: S[(j-1)*nS/NS] < x < S[j*nS/NS]  (or the end of S)
    uint64_t l = (uint64_t(lo - 1) * nS) / NS + 1;
    uint64_t h = (lo == NS) ? nS : (uint64_t(lo) * nS) / NS;
    while (l < h) {
        const uint64_t mid = (l + h) >> 1;
        const uint32_t y = S[mid];
        if (y == x) return 1;
        if (y < x)
            l = mid + 1;
        else
            h = mid;
    }
    return 0;
}

// ------------------------------------------------------ decoupled look-back
// state[s] of segment s of a materialising launch: published flag (bit 63) |
// inclusive flag (bit 62) | count (bits 0-32; a pair's intersection is below
// 2^33).  0 means "not published"; the launch's last workgroup clears the
// array, so launches need no memset (and replay correctly from a captured
// graph).
constexpr uint64_t kLbCountMask = (1ull << 33) - 1;
constexpr uint64_t kLbPublished = 1ull << 63;
constexpr uint64_t kLbInclusive = 1ull << 62;

// Called by all 64 lanes of one wave of segment s's workgroup: publishes the
// segment's count, adds up its predecessors' counts 64 states per round
// (lane l reads state[j - l]; the closest inclusive state ends the walk once
// every state up to it is published; before segment 0 there is an implicit
// inclusive 0), publishes the segment's inclusive prefix and returns the
// exclusive one.  Segments must be numbered in the order they start (a
// ticket), so that every predecessor is running or done.  The counts are the
// whole payload: relaxed atomics suffice, and polling without acquire keeps
// the CU's L1 intact for the workgroups still working beside this one.
__device__ __forceinline__ uint64_t lookback_exclusive(uint64_t* __restrict__ state, uint32_t s, uint64_t total) {
    const uint32_t lane = lane_id();
    if (lane == 0) __hip_atomic_store(&state[s], kLbPublished | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t pre = 0;
    int64_t j = int64_t(s) - 1;
    while (j >= 0) {
        const int64_t idx = j - int64_t(lane);
        const uint64_t v = idx >= 0 ? __hip_atomic_load(&state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : (kLbPublished | kLbInclusive);
        const uint64_t incl = __ballot((v & kLbInclusive) != 0);
        const uint64_t pub = __ballot((v & kLbPublished) != 0);
        // lanes [0, last] must be published: up to the closest inclusive, else all 64
        const uint32_t last = incl ? uint32_t(__builtin_ctzll(incl)) : WAVE - 1;
        const uint64_t need = last == WAVE - 1 ? ~0ull : ((2ull << last) - 1);
        if ((pub & need) != need) {
            __builtin_amdgcn_s_sleep(16);
            continue;
        }
        pre += wave_sum_u64(lane <= last ? (v & kLbCountMask) : 0ull);
        if (incl) break;
        j -= WAVE;
    }
    if (lane == 0)
        __hip_atomic_store(&state[s], kLbPublished | kLbInclusive | (pre + total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    return pre;
}

// sample_search for R keys at once (lanes with valid[r] false report no
// match): the phases run in lockstep over the keys, so each step issues R
// independent LDS or global loads.  Bit r of the result: x[r] is in S.
template <int NS, int R>
__device__ __forceinline__ uint32_t sample_search_multi(const uint32_t* __restrict__ S, uint64_t nS,
                                                        const uint32_t* cache, const uint32_t (&x)[R],
                                                        const bool (&valid)[R]) {
    static_assert((NS & (NS - 1)) == 0, "NS must be a power of two");
    uint32_t lo[R];
#pragma unroll
    for (int r = 0; r < R; ++r) lo[r] = 0;
    // phase 1: lo = number of samples <= x (branch-free over a power-of-two NS)
    constexpr int LOG2NS = __builtin_ctz(NS);
#pragma unroll
    for (int k = LOG2NS - 1; k >= 0; --k) {
        const uint32_t step = 1u << k;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (cache[lo[r] + step - 1] <= x[r]) lo[r] += step;
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (lo[r] < NS && cache[lo[r]] <= x[r]) lo[r] += 1;
    // phase 2: the lower bound of x in S[l, h), where S[(j-1)*nS/NS] < x <
    // S[j*nS/NS] (or the end of S), by a branch-free search of a fixed
    // number of steps for every key: 32-bit positions within the interval
    // (its width is at most ceil(nS / NS)), one load per key and step, the
    // loads of the R keys independent
    uint32_t hit = 0;
    const uint32_t* B[R];
    uint32_t n[R], pos[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        B[r] = S;
        n[r] = 0;
        pos[r] = 0;
        if (!valid[r] || lo[r] == 0) continue;  // x < S[0]
        if (cache[lo[r] - 1] == x[r]) {
            hit |= 1u << r;
            continue;
        }
        const uint64_t l = (uint64_t(lo[r] - 1) * nS) / NS + 1;
        const uint64_t h = (lo[r] == NS) ? nS : (uint64_t(lo[r]) * nS) / NS;
        if (l < h) {
            B[r] = S + l;
            n[r] = uint32_t(h - l);
        }
    }
    const uint64_t W = (nS + NS - 1) / NS;  // widest interval
    const uint32_t T = W ? 64u - uint32_t(__builtin_clzll(W)) : 0u;  // 2^T > W
    for (uint32_t k = T; k-- > 0;) {
        const uint32_t step = 1u << k;
        uint32_t y[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // position pos + step - 1, clamped into [0, n) (a key with n = 0 reads B[0] = S[0])
            const uint32_t q = pos[r] + step;
            y[r] = B[r][(q <= n[r] ? q : (n[r] ? n[r] : 1u)) - 1u];
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (pos[r] + step <= n[r] && y[r] < x[r]) pos[r] += step;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t y = B[r][pos[r] < n[r] ? pos[r] : 0u];
        if (pos[r] < n[r] && y == x[r]) hit |= 1u << r;
    }
    return hit;
}

// ------------------------------------------------------------ pair indexing
// triangular_index of include/util.hpp:68-70 in 64-bit arithmetic.
__device__ __forceinline__ uint64_t tri_index(uint64_t n, uint64_t i, uint64_t j) {
    return (n * (n - 1) / 2) - (n - i) * ((n - i) - 1) / 2 + j - i - 1;
}

// One GPU invocation ("run") of the reference drivers: tile A x tile B with
// the output layout of the run's counts buffer (ip.cu:145, helpers.cuh:16-22).
struct RunArgs {
    uint32_t aStart, aEnd, aId;
    uint32_t bStart, bEnd, bId;
    uint32_t n;  // numOfSets of the layout
};

template <bool SELF>
__device__ __forceinline__ uint64_t run_index(const RunArgs& r, uint32_t a, uint32_t b) {
    const uint64_t ia = uint64_t(a) - uint64_t(r.aId) * r.n, ib = uint64_t(b) - uint64_t(r.bId) * r.n;
    return SELF ? tri_index(r.n, ia, ib) : ia * r.n + ib;
}

// Inverse of the row-major strict-upper-triangle enumeration of an L x L
// tile: p -> (i, j), 0 <= i < j < L.
__device__ __forceinline__ void tri_decode(uint64_t p, uint64_t L, uint32_t& i, uint32_t& j) {
    // row i starts at r(i) = i*L - i*(i+1)/2
    const double Ld = double(L);
    double disc = (2.0 * Ld - 1.0) * (2.0 * Ld - 1.0) - 8.0 * double(p);
    if (disc < 0) disc = 0;
    int64_t ii = int64_t(((2.0 * Ld - 1.0) - sqrt(disc)) * 0.5);
    if (ii < 0) ii = 0;
    if (ii > int64_t(L) - 2) ii = int64_t(L) - 2;
    auto rowStart = [&](int64_t r) { return uint64_t(r) * L - uint64_t(r) * uint64_t(r + 1) / 2; };
    while (ii > 0 && rowStart(ii) > p) --ii;
    while (ii + 1 <= int64_t(L) - 2 && rowStart(ii + 1) <= p) ++ii;
    i = uint32_t(ii);
    j = uint32_t(p - rowStart(ii) + uint64_t(ii) + 1);
}

// Sets bit x of the 32-bit-word bitmap bm for the elements x of the sorted
// unique list S[c*64, c*64+64) handled by this wave (chunk c; x is this lane's
// element, valid when c*64 + lane < n).  Elements of one word are
// consecutive, so lanes that share a word OR their bits together (segmented
// OR over the wave with shuffles) and the segment's first lane writes the
// word: a plain store inside the chunk, an atomic OR only for the chunk's
// first and last word, which the neighbouring chunks may share.  bm must be
// zero beforehand.
__device__ __forceinline__ void or_bits_chunk(uint32_t x, bool valid, uint32_t* __restrict__ bm) {
    const uint32_t lane = lane_id();
    const uint32_t w = valid ? (x >> 5) : 0xFFFFFFFFu;
    uint32_t bits = valid ? (1u << (x & 31)) : 0u;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t vb = __shfl_down(bits, off);
        const uint32_t vw = __shfl_down(w, off);
        if (lane + off < WAVE && vw == w) bits |= vb;
    }
    const uint32_t wprev = __shfl_up(w, 1);
    const uint32_t w0 = __shfl(w, 0), wlast = __shfl(w, WAVE - 1);
    if (valid && (lane == 0 || wprev != w)) {
        if (w == w0 || w == wlast)
            atomicOr(&bm[w], bits);
        else
            bm[w] = bits;
    }
}

// Chunks cb, cb + stride, ... (UNROLL of them) of S[0, n): all loads are
// issued before the first chunk is combined, so a wave keeps UNROLL chunk
// loads in flight.
constexpr int BUILD_UNROLL = 8;

// Bits are set at x - base (a bitmap over [base, base + 32 * words)).
__device__ __forceinline__ void or_bits_chunks(const uint32_t* __restrict__ S, uint64_t n, uint64_t cb,
                                               uint64_t stride, uint32_t* __restrict__ bm, uint32_t base = 0) {
    const uint32_t lane = lane_id();
    uint32_t x[BUILD_UNROLL];
    bool v[BUILD_UNROLL];
#pragma unroll
    for (int u = 0; u < BUILD_UNROLL; ++u) {
        const uint64_t i = (cb + uint64_t(u) * stride) * WAVE + lane;
        v[u] = i < n;
        x[u] = v[u] ? S[i] - base : 0u;
    }
#pragma unroll
    for (int u = 0; u < BUILD_UNROLL; ++u)
        if (cb + uint64_t(u) * stride < (n + WAVE - 1) / WAVE) or_bits_chunk(x[u], v[u], bm);
}

}  // namespace si_dev
