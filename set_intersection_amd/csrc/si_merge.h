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
    static constexpr int CPT = (T / 4 + 4 + NT - 1) / NT;  // 16-byte chunks per thread per refill
    static_assert((T & (T - 1)) == 0, "tile must be a power of two");
};

template <int NT, int VT>
struct MergeSmem {
    uint32_t wa[NT * VT];
    uint32_t wb[NT * VT];
    uint32_t ends[NT];
    uint32_t wave_tot[NT / WAVE];
    uint64_t red[NT / WAVE];
    uint64_t seg[4];
};

// Loads of one window refill, held in registers across the tile's merge.
template <int NT, int VT>
struct Refill {
    using C = MergeCfg<NT, VT>;
    uint4 v[C::CPT];
    uint64_t cA0, cB0;     // first chunk index of the A / B ranges
    uint32_t nCA, nCB;     // chunk counts
    uint64_t gA0, gA1, gB0, gB1;  // element ranges [g0, g1) to write

    __device__ __forceinline__ void issue(const uint32_t* __restrict__ base, uint64_t a0, uint64_t a1, uint64_t b0,
                                          uint64_t b1) {
        gA0 = a0;
        gA1 = a1;
        gB0 = b0;
        gB1 = b1;
        cA0 = a0 >> 2;
        nCA = a1 > a0 ? uint32_t(((a1 + 3) >> 2) - cA0) : 0u;
        cB0 = b0 >> 2;
        nCB = b1 > b0 ? uint32_t(((b1 + 3) >> 2) - cB0) : 0u;
        const uint4* __restrict__ b4 = reinterpret_cast<const uint4*>(base);
#pragma unroll
        for (int j = 0; j < C::CPT; ++j) {
            const uint32_t c = threadIdx.x + uint32_t(j) * NT;
            if (c < nCA)
                v[j] = b4[cA0 + c];
            else if (c < nCA + nCB)
                v[j] = b4[cB0 + (c - nCA)];
        }
    }

    __device__ __forceinline__ static void put(uint32_t* w, const uint4& x, uint64_t g, uint64_t lo, uint64_t hi) {
        if (g >= lo && g + 4 <= hi) {
            *reinterpret_cast<uint4*>(&w[uint32_t(g) & C::M]) = x;
        } else {
            if (g + 0 >= lo && g + 0 < hi) w[uint32_t(g + 0) & C::M] = x.x;
            if (g + 1 >= lo && g + 1 < hi) w[uint32_t(g + 1) & C::M] = x.y;
            if (g + 2 >= lo && g + 2 < hi) w[uint32_t(g + 2) & C::M] = x.z;
            if (g + 3 >= lo && g + 3 < hi) w[uint32_t(g + 3) & C::M] = x.w;
        }
    }

    __device__ __forceinline__ void commit(MergeSmem<NT, VT>& sm) const {
#pragma unroll
        for (int j = 0; j < C::CPT; ++j) {
            const uint32_t c = threadIdx.x + uint32_t(j) * NT;
            if (c < nCA)
                put(sm.wa, v[j], (cA0 + c) * 4, gA0, gA1);
            else if (c < nCA + nCB)
                put(sm.wb, v[j], (cB0 + (c - nCA)) * 4, gB0, gB1);
        }
    }
};

// Streams the merge path from (ia, ib) to (iaEnd, ibEnd) of sets A (global
// element offset oA, size nA) and B (oB, nB) through the LDS windows.
// Returns this thread's match count.  With MAT, matches are written in order
// to stage[0 ..) (the segment's staging slot).
template <int NT, int VT, bool MAT>
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
        const uint32_t gA = uint32_t(oA + ia);  // slot base (mod T) of the windows
        const uint32_t gB = uint32_t(oB + ib);
        auto WA = [&](uint32_t i) { return sm.wa[(gA + i) & C::M]; };
        auto WB = [&](uint32_t i) { return sm.wb[(gB + i) & C::M]; };

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
        uint32_t x = x0, y = y0;
        uint32_t av = x < x1 ? WA(x) : 0u;
        uint32_t bv = y < nBw ? WB(y) : 0u;
        const uint32_t steps = (x1 - x0) + (y1 - y0);
        uint32_t cnt = 0;
        uint32_t mval[MAT ? VT : 1];
        uint32_t mbits = 0;
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
            uint64_t pos = written + before + excl;
#pragma unroll
            for (int k = 0; k < VT; ++k) {
                if (mbits & (1u << k)) stage[pos++] = mval[k];
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
