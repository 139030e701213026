// Workgroup-level streaming merge path, shared by the merge-path kernels
// (k_merge.hip) and the batched dispatcher (k_batched.hip).
//
// A workgroup of NT threads streams the merge path of sets A and B from
// (ia, ib) to (iaEnd, ibEnd) in tiles of T = NT*VT merged elements through two
// circular LDS windows (T elements of A, T of B).  After each tile only the
// consumed elements are re-fetched, with aligned 128-bit global loads issued
// before the tile's merge and committed to LDS after it.  Each thread finds
// its split by binary search in LDS at the end of its VT-element diagonal
// range (its start is its neighbour's end).  Ties put the A element first, so
// an A element is matched against the B element next in line when it is
// merged; no partition can split a matching pair.
#pragma once

#include "si_device.h"

namespace si_dev {

template <int NT, int VT>
struct MergeCfg {
    static constexpr int T = NT * VT;          // merged elements per tile
    static constexpr uint32_t M = T - 1;       // circular window mask
    static constexpr int B0 = 2 * T;           // start of the B window in MergeSmem::w
    static constexpr int CPT = (T / 4 + 4 + NT - 1) / NT;  // 16-byte chunks per thread per refill
    static_assert((T & (T - 1)) == 0, "tile must be a power of two");
};

template <int NT, int VT>
struct MergeSmem {
    // Circular windows of T elements each, stored twice (slot s and s + T) so
    // that T consecutive elements from any start slot are contiguous: A in
    // w[0, 2T), B in w[2T, 4T).
    uint32_t w[4 * NT * VT];
    uint32_t ends[NT];
    uint32_t wave_tot[NT / WAVE];
    uint64_t red[NT / WAVE];
    uint64_t seg[4];
};

// Loads of one window refill, held in registers across the tile's merge.
// Element ranges are absolute indices into the element array; per-lane work
// uses 32-bit offsets from uniform (scalar) base pointers.
template <int NT, int VT>
struct Refill {
    using C = MergeCfg<NT, VT>;
    uint4 v[C::CPT];
    const uint4* pA;
    const uint4* pB;
    uint32_t nCA, nCB;        // chunk counts
    uint32_t phA, phB;        // offset of the first needed element in its chunk
    uint32_t lenA, lenB;      // needed elements
    uint32_t sA, sB;          // LDS slot of the first chunk's first element

    __device__ __forceinline__ void issue(const uint32_t* __restrict__ base, uint64_t a0, uint64_t a1, uint64_t b0,
                                          uint64_t b1) {
        const uint4* b4 = reinterpret_cast<const uint4*>(base);
        lenA = a1 > a0 ? uint32_t(a1 - a0) : 0u;
        lenB = b1 > b0 ? uint32_t(b1 - b0) : 0u;
        nCA = lenA ? uint32_t(((a1 + 3) >> 2) - (a0 >> 2)) : 0u;
        nCB = lenB ? uint32_t(((b1 + 3) >> 2) - (b0 >> 2)) : 0u;
        phA = uint32_t(a0 & 3);
        phB = uint32_t(b0 & 3);
        sA = uint32_t(a0 & ~3ull) & C::M;
        sB = uint32_t(b0 & ~3ull) & C::M;
        pA = b4 + (a0 >> 2);
        pB = b4 + (b0 >> 2);
        // every lane issues all CPT loads back to back: out-of-range chunk
        // numbers are clamped to the last needed chunk (in bounds, not committed)
        const uint32_t tot = nCA + nCB;
        if (tot) {
#pragma unroll
            for (int j = 0; j < C::CPT; ++j) {
                uint32_t c = threadIdx.x + uint32_t(j) * NT;
                c = c < tot ? c : tot - 1;
                const uint4* p = c < nCA ? pA + c : pB + (c - nCA);
                v[j] = *p;
            }
        }
    }

    // chunk c of a range: elements rel = 4c - ph + q, valid for 0 <= rel < len;
    // written to slot and to its mirror slot + T
    __device__ __forceinline__ static void put(uint32_t* w, const uint4& x, uint32_t c, uint32_t ph, uint32_t len,
                                               uint32_t s0) {
        const int rel = int(4 * c) - int(ph);
        const uint32_t slot = (s0 + 4 * c) & C::M;
        if (rel >= 0 && rel + 4 <= int(len)) {
            *reinterpret_cast<uint4*>(&w[slot]) = x;
            *reinterpret_cast<uint4*>(&w[slot + C::T]) = x;
        } else {
            const uint32_t e[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (rel + q >= 0 && rel + q < int(len)) {
                    w[slot + q] = e[q];
                    w[slot + q + C::T] = e[q];
                }
        }
    }

    __device__ __forceinline__ void commit(MergeSmem<NT, VT>& sm) const {
#pragma unroll
        for (int j = 0; j < C::CPT; ++j) {
            const uint32_t c = threadIdx.x + uint32_t(j) * NT;
            if (c < nCA)
                put(sm.w, v[j], c, phA, lenA, sA);
            else if (c < nCA + nCB)
                put(sm.w + C::B0, v[j], c - nCA, phB, lenB, sB);
        }
    }
};

// Streams the merge path from (ia, ib) to (iaEnd, ibEnd) of sets A (global
// element offset oA, size nA) and B (oB, nB) through the LDS windows.
// Returns this thread's match count.  With MAT, matches are written in order
// to stage[0 ..) (the segment's staging slot).
template <int NT, int VT, bool MAT, bool DUAL = false>
__device__ uint64_t merge_segment(const uint32_t* __restrict__ base, uint64_t oA, uint64_t nA, uint64_t oB,
                                  uint64_t nB, uint64_t ia, uint64_t ib, uint64_t iaEnd, uint64_t ibEnd,
                                  MergeSmem<NT, VT>& sm, uint32_t* __restrict__ stage) {
    using C = MergeCfg<NT, VT>;
    constexpr uint32_t T = C::T;
    const uint32_t tid = threadIdx.x;
    uint64_t myCount = 0;
    uint64_t written = 0;  // MAT: matches already staged by this workgroup

    uint64_t d = ia + ib;
    const uint64_t dEnd = iaEnd + ibEnd;
    if (d >= dEnd) return 0;

    // initial fill: A[ia, ia+T) and B[ib, ib+T), clipped to the sets
    {
        Refill<NT, VT> r;
        const uint64_t a1 = (ia + T < nA ? ia + T : nA);
        r.issue(base, oA + ia, oA + a1, 0, 0);
        r.commit(sm);
        const uint64_t b1 = (ib + T < nB ? ib + T : nB);
        r.issue(base, 0, 0, oB + ib, oB + b1);
        r.commit(sm);
    }
    __syncthreads();

    while (d < dEnd) {
        const uint32_t Tc = uint32_t((dEnd - d) < T ? (dEnd - d) : T);
        const uint32_t nAw = uint32_t((nA - ia) < T ? (nA - ia) : T);
        const uint32_t nBw = uint32_t((nB - ib) < T ? (nB - ib) : T);
        // window bases: A[ia + i] is at wA[i], B[ib + i] at wB[i], for i <= T
        const uint32_t* wA = sm.w + (uint32_t(oA + ia) & C::M);
        const uint32_t* wB = sm.w + C::B0 + (uint32_t(oB + ib) & C::M);
        auto WA = [&](uint32_t i) { return wA[i]; };
        auto WB = [&](uint32_t i) { return wB[i]; };
        // interior tile: full windows and a full tile need no bounds checks
        const bool full = (Tc == T) && (nAw == T) && (nBw == T);

        // 1. split at the end of this thread's diagonal range
        const uint32_t dd = ((tid + 1) * VT < Tc) ? (tid + 1) * VT : Tc;
        {
            uint32_t lo = dd > nBw ? dd - nBw : 0u;
            uint32_t hi = dd < nAw ? dd : nAw;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (WA(mid) <= WB(dd - 1 - mid))
                    lo = mid + 1;
                else
                    hi = mid;
            }
            sm.ends[tid] = lo;
        }
        __syncthreads();
        const uint32_t ds = (tid * VT < Tc) ? tid * VT : Tc;
        const uint32_t x0 = tid ? sm.ends[tid - 1] : 0u;
        const uint32_t x1 = sm.ends[tid];
        const uint32_t y0 = ds - x0;
        const uint32_t y1 = dd - x1;
        const uint32_t ca = sm.ends[NT - 1];
        const uint32_t cb = Tc - ca;

        // 2. fetch the elements that replace the consumed ones (registers)
        Refill<NT, VT> r;
        {
            const uint64_t a0 = ia + T < nA ? ia + T : nA;
            const uint64_t a1 = ia + T + ca < nA ? ia + T + ca : nA;
            const uint64_t b0 = ib + T < nB ? ib + T : nB;
            const uint64_t b1 = ib + T + cb < nB ? ib + T + cb : nB;
            r.issue(base, oA + a0, oA + a1, oB + b0, oB + b1);
        }

        // 3. merge this thread's VT steps; A is taken on ties and matched
        //    against the current B head (which may lie past y1).
        uint32_t cnt = 0;
        uint32_t mval[MAT ? VT : 1];
        uint32_t mbits = 0;
        if (full) {
            // Every thread has exactly VT steps and the true heads A[x], B[y]
            // are in the windows for x <= x1, y <= y1 (x1 == T or y1 == T only
            // when the other side has no steps), so the merge-path order alone
            // decides: take A iff A[x] <= B[y]; a tie is a match.
            if (DUAL) {
                // both heads re-read every step: y = (x0 + y0 + k) - x, so the
                // B address is a fixed base minus x plus the step constant
                uint32_t x = x0;
                const uint32_t* pB = wB + (x0 + y0);
                uint32_t av = wA[x], bv = pB[-int(x)];
#pragma unroll
                for (int k = 0; k < VT; ++k) {
                    const bool takeA = av <= bv;
                    const bool hit = av == bv;
                    cnt += hit;
                    if (MAT) {
                        mval[k] = av;
                        mbits |= uint32_t(hit) << k;
                    }
                    x += takeA;
                    av = wA[x];
                    bv = pB[k + 1 - int(x)];
                }
            } else {
                const uint32_t* pa = wA + x0;
                const uint32_t* pb = wB + y0;
                uint32_t av = *pa, bv = *pb;
#pragma unroll
                for (int k = 0; k < VT; ++k) {
                    const bool takeA = av <= bv;
                    const bool hit = av == bv;
                    cnt += hit;
                    if (MAT) {
                        mval[k] = av;
                        mbits |= uint32_t(hit) << k;
                    }
                    pa += takeA;
                    pb += !takeA;
                    const uint32_t nv = *(takeA ? pa : pb);
                    av = takeA ? nv : av;
                    bv = takeA ? bv : nv;
                }
            }
        } else {
            uint32_t x = x0, y = y0;
            uint32_t av = x < x1 ? WA(x) : 0u;
            uint32_t bv = y < nBw ? WB(y) : 0u;
            const uint32_t steps = (x1 - x0) + (y1 - y0);
#pragma unroll
            for (int k = 0; k < VT; ++k) {
                if (uint32_t(k) < steps) {
                    const bool takeA = (x < x1) && (y >= y1 || av <= bv);
                    const bool hit = takeA && (y < nBw) && (av == bv);
                    cnt += hit;
                    if (MAT) {
                        mval[k] = av;
                        mbits |= uint32_t(hit) << k;
                    }
                    if (takeA) {
                        ++x;
                        if (x < x1) av = WA(x);
                    } else {
                        ++y;
                        if (y < nBw) bv = WB(y);
                    }
                }
            }
        }
        myCount += cnt;

        uint32_t excl = 0, wtot = 0;
        if (MAT) {
            const uint32_t incl = wave_inclusive_scan_u32(cnt);
            excl = incl - cnt;
            if (lane_id() == WAVE - 1) sm.wave_tot[tid / WAVE] = incl;
        }
        __syncthreads();  // all window reads of this tile are done

        // 4. commit the refill; with MAT, stage this tile's matches
        r.commit(sm);
        if (MAT) {
            const uint32_t w = tid / WAVE;
            uint32_t before = 0;
#pragma unroll
            for (int i = 0; i < NT / WAVE; ++i) {
                const uint32_t t = sm.wave_tot[i];
                if (i < int(w)) before += t;
                wtot += t;
            }
            uint32_t* dst = stage + written + before + excl;
#pragma unroll
            for (int k = 0; k < VT; ++k) {
                if (mbits & (1u << k)) *dst++ = mval[k];
            }
            written += wtot;
        }
        ia += ca;
        ib += cb;
        d += Tc;
        __syncthreads();
    }
    return myCount;
}

// --------------------------------------------------------------------------
// Prefetching variant for long segments (SISI).  Each side keeps a ring of
// R = 2T elements in LDS: while tile k is merged, the ring already holds the
// next T elements past tile k's window, and the loads that replace tile k's
// consumed elements are issued during tile k but committed only at the end
// of tile k+1 (two register sets), so every global load has a whole extra
// tile of time to arrive.
//   ring content at the start of tile k: A[ia_k, ia_k + 2T) (clipped to |A|)
//   tile k issues A[ia_k + 2T, ia_{k+1} + 2T) (= ca_k elements)
//   tile k commits tile k-1's loads into the slots consumed by tile k-1.
template <int NT, int VT>
struct RingCfg {
    static constexpr int T = NT * VT;
    static constexpr uint32_t R = 2 * T;      // ring capacity per side
    static constexpr uint32_t M = R - 1;
    static constexpr int CPT = (T / 4 + 4 + NT - 1) / NT;
    static_assert((T & (T - 1)) == 0, "tile must be a power of two");
};

template <int NT, int VT>
struct RingSmem {
    uint32_t ring[2 * 2 * NT * VT];  // A ring [0, R), B ring [R, 2R)
    uint32_t ends[NT];
    uint32_t wave_tot[NT / WAVE];
    uint64_t red[NT / WAVE];
    uint64_t seg[4];
};

// A refill of at most T elements per side, held in registers.
template <int NT, int VT>
struct RingLoad {
    using C = RingCfg<NT, VT>;
    uint4 v[C::CPT];
    uint32_t nCA, nCB, phA, phB, lenA, lenB, sA, sB;

    __device__ __forceinline__ void none() { nCA = nCB = 0; }

    __device__ __forceinline__ void issue(const uint32_t* __restrict__ base, uint64_t a0, uint64_t a1, uint64_t b0,
                                          uint64_t b1) {
        const uint4* b4 = reinterpret_cast<const uint4*>(base);
        lenA = a1 > a0 ? uint32_t(a1 - a0) : 0u;
        lenB = b1 > b0 ? uint32_t(b1 - b0) : 0u;
        nCA = lenA ? uint32_t(((a1 + 3) >> 2) - (a0 >> 2)) : 0u;
        nCB = lenB ? uint32_t(((b1 + 3) >> 2) - (b0 >> 2)) : 0u;
        phA = uint32_t(a0 & 3);
        phB = uint32_t(b0 & 3);
        sA = uint32_t(a0 & ~3ull) & C::M;
        sB = uint32_t(b0 & ~3ull) & C::M;
        const uint4* pA = b4 + (a0 >> 2);
        const uint4* pB = b4 + (b0 >> 2);
        const uint32_t tot = nCA + nCB;
        if (tot) {
#pragma unroll
            for (int j = 0; j < C::CPT; ++j) {
                uint32_t c = threadIdx.x + uint32_t(j) * NT;
                c = c < tot ? c : tot - 1;
                const uint4* p = c < nCA ? pA + c : pB + (c - nCA);
                v[j] = *p;
            }
        }
    }

    __device__ __forceinline__ static void put(uint32_t* w, const uint4& x, uint32_t c, uint32_t ph, uint32_t len,
                                               uint32_t s0) {
        const int rel = int(4 * c) - int(ph);
        const uint32_t slot = (s0 + 4 * c) & C::M;
        if (rel >= 0 && rel + 4 <= int(len)) {
            *reinterpret_cast<uint4*>(&w[slot]) = x;
        } else {
            const uint32_t e[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (rel + q >= 0 && rel + q < int(len)) w[slot + q] = e[q];
        }
    }

    __device__ __forceinline__ void commit(RingSmem<NT, VT>& sm) const {
#pragma unroll
        for (int j = 0; j < C::CPT; ++j) {
            const uint32_t c = threadIdx.x + uint32_t(j) * NT;
            if (c < nCA)
                put(sm.ring, v[j], c, phA, lenA, sA);
            else if (c < nCA + nCB)
                put(sm.ring + C::R, v[j], c - nCA, phB, lenB, sB);
        }
    }
};

template <int NT, int VT, bool MAT, bool DUAL>
__device__ uint64_t merge_segment_pf(const uint32_t* __restrict__ base, uint64_t oA, uint64_t nA, uint64_t oB,
                                     uint64_t nB, uint64_t ia, uint64_t ib, uint64_t iaEnd, uint64_t ibEnd,
                                     RingSmem<NT, VT>& sm, uint32_t* __restrict__ stage) {
    using C = RingCfg<NT, VT>;
    constexpr uint32_t T = C::T;
    const uint32_t tid = threadIdx.x;
    uint64_t myCount = 0;
    uint64_t written = 0;

    uint64_t d = ia + ib;
    const uint64_t dEnd = iaEnd + ibEnd;
    if (d >= dEnd) return 0;

    // initial fill: A[ia, ia + 2T) and B[ib, ib + 2T), in four parts of <= T
    {
        RingLoad<NT, VT> r;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint64_t a0 = ia + uint64_t(h) * T < nA ? ia + uint64_t(h) * T : nA;
            const uint64_t a1 = ia + uint64_t(h + 1) * T < nA ? ia + uint64_t(h + 1) * T : nA;
            r.issue(base, oA + a0, oA + a1, 0, 0);
            r.commit(sm);
            const uint64_t b0 = ib + uint64_t(h) * T < nB ? ib + uint64_t(h) * T : nB;
            const uint64_t b1 = ib + uint64_t(h + 1) * T < nB ? ib + uint64_t(h + 1) * T : nB;
            r.issue(base, 0, 0, oB + b0, oB + b1);
            r.commit(sm);
        }
    }
    __syncthreads();

    RingLoad<NT, VT> pend;
    pend.none();
    while (d < dEnd) {
        const uint32_t Tc = uint32_t((dEnd - d) < T ? (dEnd - d) : T);
        const uint32_t nAw = uint32_t((nA - ia) < T ? (nA - ia) : T);
        const uint32_t nBw = uint32_t((nB - ib) < T ? (nB - ib) : T);
        const uint32_t sa = uint32_t(oA + ia) & C::M;  // ring slot of A[ia]
        const uint32_t sb = uint32_t(oB + ib) & C::M;
        const uint32_t* rA = sm.ring;
        const uint32_t* rB = sm.ring + C::R;
        auto WA = [&](uint32_t i) { return rA[(sa + i) & C::M]; };
        auto WB = [&](uint32_t i) { return rB[(sb + i) & C::M]; };
        const bool full = (Tc == T) && (nAw == T) && (nBw == T);

        // 1. split at the end of this thread's diagonal range
        const uint32_t dd = ((tid + 1) * VT < Tc) ? (tid + 1) * VT : Tc;
        {
            uint32_t lo = dd > nBw ? dd - nBw : 0u;
            uint32_t hi = dd < nAw ? dd : nAw;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (WA(mid) <= WB(dd - 1 - mid))
                    lo = mid + 1;
                else
                    hi = mid;
            }
            sm.ends[tid] = lo;
        }
        __syncthreads();
        const uint32_t ds = (tid * VT < Tc) ? tid * VT : Tc;
        const uint32_t x0 = tid ? sm.ends[tid - 1] : 0u;
        const uint32_t x1 = sm.ends[tid];
        const uint32_t y0 = ds - x0;
        const uint32_t y1 = dd - x1;
        const uint32_t ca = sm.ends[NT - 1];
        const uint32_t cb = Tc - ca;

        // 2. issue the loads that replace this tile's consumed elements, 2T ahead
        RingLoad<NT, VT> nxt;
        {
            const uint64_t a0 = ia + 2 * T < nA ? ia + 2 * T : nA;
            const uint64_t a1 = ia + 2 * T + ca < nA ? ia + 2 * T + ca : nA;
            const uint64_t b0 = ib + 2 * T < nB ? ib + 2 * T : nB;
            const uint64_t b1 = ib + 2 * T + cb < nB ? ib + 2 * T + cb : nB;
            nxt.issue(base, oA + a0, oA + a1, oB + b0, oB + b1);
        }

        // 3. merge this thread's VT steps (A first on ties; a tie is a match)
        uint32_t cnt = 0;
        uint32_t mval[MAT ? VT : 1];
        uint32_t mbits = 0;
        if (full) {
            if (DUAL) {
                uint32_t x = x0;
                const uint32_t cbase = sb + x0 + y0;  // B slot of step k: (cbase + k - x) & M
                uint32_t av = rA[(sa + x) & C::M], bv = rB[(cbase - x) & C::M];
#pragma unroll
                for (int k = 0; k < VT; ++k) {
                    const bool takeA = av <= bv;
                    const bool hit = av == bv;
                    cnt += hit;
                    if (MAT) {
                        mval[k] = av;
                        mbits |= uint32_t(hit) << k;
                    }
                    x += takeA;
                    av = rA[(sa + x) & C::M];
                    bv = rB[(cbase + k + 1 - x) & C::M];
                }
            } else {
                uint32_t pa = sa + x0, pb = C::R + sb + y0;  // unmasked ring positions
                uint32_t av = WA(x0), bv = WB(y0);
#pragma unroll
                for (int k = 0; k < VT; ++k) {
                    const bool takeA = av <= bv;
                    const bool hit = av == bv;
                    cnt += hit;
                    if (MAT) {
                        mval[k] = av;
                        mbits |= uint32_t(hit) << k;
                    }
                    pa += takeA;
                    pb += !takeA;
                    const uint32_t slot = takeA ? (pa & C::M) : (C::R + ((pb - C::R) & C::M));
                    const uint32_t nv = sm.ring[slot];
                    av = takeA ? nv : av;
                    bv = takeA ? bv : nv;
                }
            }
        } else {
            uint32_t x = x0, y = y0;
            uint32_t av = x < x1 ? WA(x) : 0u;
            uint32_t bv = y < nBw ? WB(y) : 0u;
            const uint32_t steps = (x1 - x0) + (y1 - y0);
#pragma unroll
            for (int k = 0; k < VT; ++k) {
                if (uint32_t(k) < steps) {
                    const bool takeA = (x < x1) && (y >= y1 || av <= bv);
                    const bool hit = takeA && (y < nBw) && (av == bv);
                    cnt += hit;
                    if (MAT) {
                        mval[k] = av;
                        mbits |= uint32_t(hit) << k;
                    }
                    if (takeA) {
                        ++x;
                        if (x < x1) av = WA(x);
                    } else {
                        ++y;
                        if (y < nBw) bv = WB(y);
                    }
                }
            }
        }
        myCount += cnt;

        uint32_t excl = 0, wtot = 0;
        if (MAT) {
            const uint32_t incl = wave_inclusive_scan_u32(cnt);
            excl = incl - cnt;
            if (lane_id() == WAVE - 1) sm.wave_tot[tid / WAVE] = incl;
        }
        __syncthreads();  // all ring reads of this tile are done

        // 4. commit the previous tile's loads; with MAT, stage this tile's matches
        pend.commit(sm);
        pend = nxt;
        if (MAT) {
            const uint32_t w = tid / WAVE;
            uint32_t before = 0;
#pragma unroll
            for (int i = 0; i < NT / WAVE; ++i) {
                const uint32_t t = sm.wave_tot[i];
                if (i < int(w)) before += t;
                wtot += t;
            }
            uint32_t* dst = stage + written + before + excl;
#pragma unroll
            for (int k = 0; k < VT; ++k) {
                if (mbits & (1u << k)) *dst++ = mval[k];
            }
            written += wtot;
        }
        ia += ca;
        ib += cb;
        d += Tc;
        __syncthreads();
    }
    return myCount;
}

}  // namespace si_dev
