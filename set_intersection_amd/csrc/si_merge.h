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
// Half-window streaming merge (SISI).  A tile of up to T = NT*VT merged
// elements consumes about T/2 from each side, so each side keeps only a window
// of W elements (a power of two a little above T/2).  Element e (absolute
// index in the dataset) lives at slot e & (W-1); the window of A holds the
// elements from A[ia] up to a multiple of 4, at least W-3 of them while A
// goes on, so every refill is a run of whole, aligned 16-byte chunks whose
// slots hold consumed elements only (no per-element masks).  The first PAD
// slots are mirrored past the end so that a thread's VT+1 consecutive reads
// need no masking.  A thread's split is feasible when it consumes at most
// capA elements of A and capB of B (W-4 while more than W-4 remain, else all
// that remain); feasibility is monotone along the diagonal, so the feasible
// threads form a prefix and the tile is cut after them.  For random data a
// tile is cut only when one side runs more than (W - 4 - T/2) ahead.
template <int NT, int VT, int WN>
struct StreamCfg {
    static constexpr uint32_t T = uint32_t(NT) * VT;
    static constexpr uint32_t W = WN;
    static constexpr uint32_t M = W - 1;
    static constexpr uint32_t PAD = (uint32_t(VT) + 2 + 3) & ~3u;  // mirrored head slots
    static constexpr uint32_t SPAN = W + PAD;                       // LDS words per side
    static_assert((W & (W - 1)) == 0, "window must be a power of two");
    static_assert(uint32_t(VT) + 1 < W, "window shorter than a thread's range");
    static_assert(VT <= 32, "match bits are kept in one 32-bit word");
};

// hits: a tile's matches compacted in order before one coalesced store (when
// they fit; a materialising kernel only)
template <int NT, int VT, int WN, int HB = 1>
struct StreamSmem {
    alignas(16) uint32_t win[2 * StreamCfg<NT, VT, WN>::SPAN];  // A window [0, SPAN), B window [SPAN, 2 SPAN)
    uint32_t hits[HB];
    uint32_t ends[NT];
    uint32_t wave_feas[NT / WAVE];
    uint32_t wave_tot[NT / WAVE];
    uint64_t red[NT / WAVE];
    uint64_t seg[4];
};

// One refill of whole 16-byte chunks over both sides, held in registers from
// issue to commit.  A tile takes at most W-4 elements of a side, so a side's
// refill is at most W/4 chunks: each thread owns chunks t + j*NT (j < CPS) of
// each side.  A side's chunks are read through a buffer resource that ends
// after its last chunk, so every lane issues all its loads with a constant
// offset (loads past the end return zeros and are not committed).
template <int NT, int VT, int WN>
struct StreamLoad {
    using C = StreamCfg<NT, VT, WN>;
    static constexpr int CPS = int((C::W / 4 + NT - 1) / NT);  // chunks per thread per side
    // a native 4 x u32 vector (the refill stays in registers)
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 va[CPS], vb[CPS];
    uint32_t nCA, nCB, sA, sB;  // chunk counts; window slot of each side's first chunk

    // chunks [ca0, ca1) of the elements (A side) and [cb0, cb1) (B side), as
    // absolute 16-byte chunk numbers of `base`; at most W/4 per side
    __device__ __forceinline__ void issue(const uint32_t* __restrict__ base, uint64_t ca0, uint64_t ca1, uint64_t cb0,
                                          uint64_t cb1) {
        nCA = uint32_t(ca1 - ca0);
        nCB = uint32_t(cb1 - cb0);
        sA = uint32_t(ca0 * 4u) & C::M;
        sB = uint32_t(cb0 * 4u) & C::M;
        const int vo = int(threadIdx.x * 16u);
        // word 3 = 0x00020000: raw 32-bit data on gfx9 parts
        if (nCA) {
            const __amdgpu_buffer_rsrc_t ra =
                __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(base) + ca0 * 4u, 0, int(nCA * 16u), 0x00020000);
#pragma unroll
            for (int j = 0; j < CPS; ++j) va[j] = __builtin_amdgcn_raw_buffer_load_b128(ra, vo, j * NT * 16, 0);
        }
        if (nCB) {
            const __amdgpu_buffer_rsrc_t rb =
                __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(base) + cb0 * 4u, 0, int(nCB * 16u), 0x00020000);
#pragma unroll
            for (int j = 0; j < CPS; ++j) vb[j] = __builtin_amdgcn_raw_buffer_load_b128(rb, vo, j * NT * 16, 0);
        }
    }

    template <typename Smem>
    __device__ __forceinline__ void commit(Smem& sm) const {
#pragma unroll
        for (int j = 0; j < CPS; ++j) {
            const uint32_t c = threadIdx.x + uint32_t(j) * NT;
            if (c < nCA) {
                const uint32_t slot = (sA + 4u * c) & C::M;
                *reinterpret_cast<u32x4*>(&sm.win[slot]) = va[j];
                if (slot < C::PAD) *reinterpret_cast<u32x4*>(&sm.win[C::W + slot]) = va[j];
            }
            if (c < nCB) {
                const uint32_t slot = (sB + 4u * c) & C::M;
                *reinterpret_cast<u32x4*>(&sm.win[C::SPAN + slot]) = vb[j];
                if (slot < C::PAD) *reinterpret_cast<u32x4*>(&sm.win[C::SPAN + C::W + slot]) = vb[j];
            }
        }
    }
};

// First i in [lo, hi] such that A[i] > B[dd - 1 - i] (hi if none), for
// hi - lo <= 2*STEP - 1, with STEP a power of two; A[i] is at
// wa[(sa + i) & M], B[j] at wb[(sb + j) & M].
// Branch-free power-of-two stepping on byte offsets: both reads are issued
// every step and out-of-range candidates are discarded.
template <uint32_t M, uint32_t STEP>
__device__ __forceinline__ uint32_t split_search(const uint32_t* wa, uint32_t sa, const uint32_t* wb, uint32_t sb,
                                                 uint32_t dd, uint32_t lo, uint32_t hi) {
    const char* pa = reinterpret_cast<const char*>(wa);
    const char* pb = reinterpret_cast<const char*>(wb);
    constexpr uint32_t M4 = M * 4u;
    const uint32_t am1 = (sa - 1u) * 4u;  // A[c - 1] at byte (am1 + 4c) & M4
    const uint32_t bk = (sb + dd) * 4u;   // B[dd - c] at byte (bk - 4c) & M4
    const uint32_t hi4 = hi * 4u;
    uint32_t x4 = lo * 4u;
#pragma unroll
    for (uint32_t s = STEP; s; s >>= 1) {
        const uint32_t c4 = x4 + 4u * s;
        const uint32_t av = *reinterpret_cast<const uint32_t*>(pa + ((am1 + c4) & M4));
        const uint32_t bv = *reinterpret_cast<const uint32_t*>(pb + ((bk - c4) & M4));
        const bool ok = (c4 <= hi4) & (av <= bv);
        x4 = ok ? c4 : x4;
    }
    return x4 >> 2;
}

// SEL: one LDS read per merge step (the next element of the side just
// consumed, selected by address) instead of re-reading both heads.
template <int NT, int VT, int WN, bool MAT, bool SEL, int HB>
__device__ uint64_t merge_segment_stream(const uint32_t* __restrict__ base, uint64_t oA, uint64_t nA, uint64_t oB,
                                         uint64_t nB, uint64_t ia, uint64_t ib, uint64_t iaEnd, uint64_t ibEnd,
                                         StreamSmem<NT, VT, WN, HB>& sm, uint32_t* __restrict__ stage,
                                         uint32_t prioSeed = kNoPrio) {
    using C = StreamCfg<NT, VT, WN>;
    constexpr uint32_t T = C::T, W = C::W, M = C::M;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    // a lane adds at most VT per tile: 2^25 over any segment of 2^33 elements
    uint32_t myCount = 0;
    uint64_t written = 0;

    uint64_t d = ia + ib;
    const uint64_t dEnd = iaEnd + ibEnd;
    if (d >= dEnd) return 0;

    // window ends in 16-byte chunks: the A window holds the elements
    // [oA + ia, 4 xA4); the chunk after each set's last element bounds them
    const uint64_t endA4 = (oA + nA + 3) >> 2, endB4 = (oB + nB + 3) >> 2;
    auto winEnd4 = [](uint64_t first, uint64_t end4) {
        const uint64_t e = (first + W) >> 2;
        return e < end4 ? e : end4;
    };
    uint64_t xA4 = winEnd4(oA + ia, endA4), xB4 = winEnd4(oB + ib, endB4);
    {
        // elements of the first chunk before A[ia] land in slots outside the window
        StreamLoad<NT, VT, WN> r;
        r.issue(base, (oA + ia) >> 2, xA4, 0, 0);
        r.commit(sm);
        r.issue(base, 0, 0, (oB + ib) >> 2, xB4);
        r.commit(sm);
    }
    __syncthreads();

    const uint32_t* rA = sm.win;
    const uint32_t* rB = sm.win + C::SPAN;
    // split prediction: A's share of the previous tile (16-bit fixed point);
    // after a tile in which a bounded search missed, the next kBackoff tiles
    // search the whole range
    constexpr uint32_t kBackoff = 8;
    uint32_t shareQ = 1u << 15, fullLeft = 0;
    uint32_t tile = 0;
    while (d < dEnd) {
        if (prioSeed != kNoPrio) set_wave_prio((tile + prioSeed) & 3u);
        const uint64_t rem64 = dEnd - d;
        const uint32_t rem = uint32_t(rem64 < T ? rem64 : T);
        // a side is cut when the window cannot hold all that remains of it;
        // a tile then consumes at most W-4 of its elements, so every head the
        // merge reads (index <= W-4) is a true element
        const bool aMore = (nA - ia) > W - 4;
        const bool bMore = (nB - ib) > W - 4;
        const uint32_t capA = aMore ? W - 4 : uint32_t(nA - ia);
        const uint32_t capB = bMore ? W - 4 : uint32_t(nB - ib);
        const uint32_t sa = uint32_t(oA + ia) & M;  // window slot of A[ia]
        const uint32_t sb = uint32_t(oB + ib) & M;

        // 1. feasible split at the end of this thread's diagonal range
        const uint32_t dd = (tid + 1) * VT < rem ? (tid + 1) * VT : rem;
        {
            const uint32_t lo = dd > capB ? dd - capB : 0u;
            const uint32_t hi0 = dd < capA ? dd : capA;
            bool feas = lo <= hi0;
            const uint32_t hi = feas ? hi0 : lo;
            // bounded search of +-R around the predicted split; exact unless
            // it ends on a bound of [L, H] that is not a bound of [lo, hi]
            // and the predicate there says the split lies beyond (then the
            // whole range is searched).  For uniform data the split stays
            // within about 150 of the prediction at the end of a tile.
            constexpr uint32_t R = 256;
            uint32_t r;
            bool missed = false;
            if (fullLeft == 0) {
                const uint32_t c0 = (dd * shareQ) >> 16;
                const uint32_t c = c0 < lo ? lo : (c0 > hi ? hi : c0);
                const uint32_t L = c > lo + (R - 1) ? c - (R - 1) : lo;
                const uint32_t H = c + R < hi ? c + R : hi;
                r = split_search<M, R>(rA, sa, rB, sb, dd, L, H);  // H - L <= 2R - 1
                missed = (r == L && L > lo && (rA[(sa + L - 1) & M] > rB[(sb + dd - L) & M])) ||
                         (r == H && H < hi && !(rA[(sa + H) & M] > rB[(sb + dd - 1 - H) & M]));
                if (missed) r = split_search<M, W / 2>(rA, sa, rB, sb, dd, lo, hi);  // hi - lo <= W - 1
            } else {
                r = split_search<M, W / 2>(rA, sa, rB, sb, dd, lo, hi);
            }
            // the true split may lie left of lo (more than capB from B) only
            // when B continues, or right of hi (more than capA from A) only
            // when A continues: test the predicate just outside
            if (r == lo && lo > 0 && bMore) feas = feas && !(rA[(sa + lo - 1) & M] > rB[(sb + dd - lo) & M]);
            if (r == hi && hi < dd && aMore) feas = feas && (rA[(sa + hi) & M] > rB[(sb + dd - 1 - hi) & M]);
            sm.ends[tid] = r;
            const uint64_t fb = __ballot(feas);
            const bool waveMissed = __ballot(missed) != 0;
            if (lane == 0) sm.wave_feas[tid / WAVE] = uint32_t(__popcll(fb)) | (waveMissed ? 0x10000u : 0u);
        }
        __syncthreads();
        uint32_t F = 0, anyMissed = 0;
#pragma unroll
        for (int i = 0; i < NT / WAVE; ++i) {
            F += sm.wave_feas[i] & 0xffffu;
            anyMissed |= sm.wave_feas[i] >> 16;
        }
        // the tile's shape is the same in every lane: keep it in scalar registers
        F = __builtin_amdgcn_readfirstlane(F);
        anyMissed = __builtin_amdgcn_readfirstlane(anyMissed);
        fullLeft = anyMissed ? kBackoff : (fullLeft ? fullLeft - 1 : 0u);
        const uint32_t Tc = F * VT < rem ? F * VT : rem;
        const uint32_t ca = __builtin_amdgcn_readfirstlane(sm.ends[F - 1]);
        const uint32_t cb = Tc - ca;
        const bool active = tid < F;
        const uint32_t ds = active ? (tid * VT < Tc ? tid * VT : Tc) : Tc;
        const uint32_t de = active ? dd : Tc;
        const uint32_t x0 = active ? (tid ? sm.ends[tid - 1] : 0u) : ca;
        const uint32_t x1 = active ? sm.ends[tid] : ca;
        const uint32_t y0 = ds - x0;
        const uint32_t y1 = de - x1;

        // 2. fetch the elements that replace the consumed ones (registers)
        // the segment's last tile needs no refill (nothing reads the window after it)
        const bool lastTile = d + Tc >= dEnd;
        StreamLoad<NT, VT, WN> r;
        const uint64_t nxA4 = winEnd4(oA + ia + ca, endA4), nxB4 = winEnd4(oB + ib + cb, endB4);
        if (!lastTile) r.issue(base, xA4, nxA4, xB4, nxB4);

        // 3. merge this thread's steps; A is taken on ties and matched against
        //    the current B head (which may lie past y1)
        uint32_t cnt = 0;
        // MAT: bit (VT-1-k) of mbits is set when step k is a match, of tbits
        // when step k takes an A element
        uint32_t mbits = 0, tbits = 0;
        const uint32_t steps = de - ds;
        if (aMore && bMore && steps == uint32_t(VT)) {
            // x1 <= W-4 and y1 <= W-4: every head read is a true element and
            // a thread's reads stay within VT+1 slots of its (masked) start.
            // A[x] at byte pa + 4x, B[ds + k - x] at byte pb + 4k - 4x
            const char* pa = reinterpret_cast<const char*>(rA + (((sa + x0) & M) - x0));
            const char* pb = reinterpret_cast<const char*>(rB + (((sb + y0) & M) - y0 + ds));
            uint32_t x = x0;
            uint32_t av = *reinterpret_cast<const uint32_t*>(pa + 4 * x);
            uint32_t bv = *reinterpret_cast<const uint32_t*>(pb - 4 * x);
            uint32_t hits = 0;
#pragma unroll
            for (int k = 0; k < VT; ++k) {
                const bool takeA = av <= bv;
                const bool hit = av == bv;
                if (MAT) {
                    mbits = 2 * mbits + uint32_t(hit);
                    tbits = 2 * tbits + uint32_t(takeA);
                } else {
                    hits += uint32_t(__popcll(__ballot(hit)));
                }
                x += uint32_t(takeA);
                // 24-bit multiply-add keeps the B address one instruction
                const char* qa = pa + 4 * x;
                const char* qb = pb + __mul24(int(x), -4) + 4 * (k + 1);
                if (SEL) {
                    const uint32_t nv = *reinterpret_cast<const uint32_t*>(takeA ? qa : qb);
                    av = takeA ? nv : av;
                    bv = takeA ? bv : nv;
                } else {
                    av = *reinterpret_cast<const uint32_t*>(qa);
                    bv = *reinterpret_cast<const uint32_t*>(qb);
                }
            }
            if (MAT)
                cnt = uint32_t(__popc(mbits));
            else if (lane == uint32_t(__ffsll(__ballot(1)) - 1))
                cnt = hits;
        } else if (steps) {
            const uint32_t nBw = uint32_t((nB - ib) < W ? (nB - ib) : W);
            uint32_t x = x0, y = y0;
            uint32_t av = x < x1 ? rA[(sa + x) & M] : 0u;
            uint32_t bv = y < nBw ? rB[(sb + y) & M] : 0u;
#pragma unroll
            for (int k = 0; k < VT; ++k) {
                bool hit = false, took = false;
                if (uint32_t(k) < steps) {
                    const bool takeA = (x < x1) && (y >= y1 || av <= bv);
                    hit = takeA && (y < nBw) && (av == bv);
                    took = takeA;
                    if (takeA) {
                        ++x;
                        if (x < x1) av = rA[(sa + x) & M];
                    } else {
                        ++y;
                        if (y < nBw) bv = rB[(sb + y) & M];
                    }
                }
                cnt += uint32_t(hit);
                if (MAT) {
                    mbits = 2 * mbits + uint32_t(hit);
                    tbits = 2 * tbits + uint32_t(took);
                }
            }
        }
        myCount += cnt;

        // MAT: a wave stages its matches in its own HW-entry part of hits
        // while the window still holds them (the refill commits only after
        // the next barrier), at wave-local offsets; after the barrier it
        // copies them to the slot.  A wave with more than HW matches instead
        // writes them to the slot straight from the window after the barrier,
        // and one more barrier then keeps the refill from overwriting them.
        constexpr uint32_t HW = uint32_t(HB) / (NT / WAVE);
        uint32_t excl = 0, wsum = 0;
        if (MAT) {
            // cnt <= VT: bit planes of the counts give the wave's prefix
            constexpr int kBits = 32 - __builtin_clz(uint32_t(VT));
            excl = wave_excl_scan_bits<kBits>(cnt, wsum);
            wsum = __builtin_amdgcn_readfirstlane(wsum);
            if (wsum <= HW) {
                // step k's A element is x0 + (A steps before k)
                uint32_t* hw = sm.hits + (tid / WAVE) * HW;
                uint32_t mb = mbits, pos = excl;
                while (mb) {
                    const uint32_t hb = 31u - uint32_t(__builtin_clz(mb));  // earliest remaining match
                    const uint32_t j = uint32_t(__popc(tbits >> (hb + 1)));
                    hw[pos++] = rA[(sa + x0 + j) & M];
                    mb &= ~(1u << hb);
                }
            }
            if (lane == 0) sm.wave_tot[tid / WAVE] = wsum;
        }
        __syncthreads();  // the tile's merge reads are done; wave totals are published

        // 4. with MAT, place the waves' matches in the slot
        uint32_t before = 0;
        uint32_t* dst = nullptr;
        if (MAT) {
            const uint32_t w = tid / WAVE;
            uint32_t wtot = 0, over = 0;
#pragma unroll
            for (int i = 0; i < NT / WAVE; ++i) {
                const uint32_t t = sm.wave_tot[i];
                if (i < int(w)) before += t;
                wtot += t;
                over |= uint32_t(t > HW);
            }
            // uniform in the wave (before) and in the workgroup (wtot, over)
            before = __builtin_amdgcn_readfirstlane(before);
            wtot = __builtin_amdgcn_readfirstlane(wtot);
            over = __builtin_amdgcn_readfirstlane(over);
            // the tile's part of the slot: a scalar base and 32-bit offsets
            dst = stage + written;
            if (wsum > HW) {
                uint32_t mb = mbits, pos = before + excl;
                while (mb) {
                    const uint32_t hb = 31u - uint32_t(__builtin_clz(mb));
                    const uint32_t j = uint32_t(__popc(tbits >> (hb + 1)));
                    dst[pos++] = rA[(sa + x0 + j) & M];
                    mb &= ~(1u << hb);
                }
            }
            if (over) __syncthreads();  // matches read from the window before the refill overwrites their slots
            written += wtot;
        }

        // 5. commit the refill; then (MAT) each staging wave stores its
        //    matches, whose LDS part only this wave rewrites (next tile)
        if (!lastTile) r.commit(sm);
        if (MAT && wsum <= HW && wsum) {
            // the wave's part of the slot as a buffer: scalar base, 32-bit
            // offsets (word 3 = 0x00020000: raw 32-bit data on gfx9 parts)
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc(dst + before, 0, int(wsum * 4u), 0x00020000);
            const uint32_t* hw = sm.hits + (tid / WAVE) * HW;
            for (uint32_t i = lane; i < wsum; i += WAVE)
                __builtin_amdgcn_raw_buffer_store_b32(hw[i], rs, int(i * 4u), 0, 0);
        }
        xA4 = nxA4;
        xB4 = nxB4;
        if (Tc) shareQ = uint32_t((uint64_t(ca) << 16) / Tc);
        ++tile;
        ia += ca;
        ib += cb;
        d += Tc;
        __syncthreads();
    }
    if (prioSeed != kNoPrio) set_wave_prio(0);
    return myCount;
}

}  // namespace si_dev
