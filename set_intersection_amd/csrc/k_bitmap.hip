// Bitmap kernels for gfx950: bitmap_dynamic (build a set's bitmap, probe the
// other set's elements) and bitmap_naive (bitmaps of all sets, AND + popcount
// for every pair).
//
// Reference behaviour
//   bitmap_dynamic.cu:240-276  SISI: constructBitmap of set 0 with atomicOr,
//                              singleInstanceIntersect probes set 1, per-block
//                              counts, thrust::reduce.
//   bitmap_dynamic.cu:279-319  MISI: per block and row a, a private global
//                              bitmap, one atomicAdd per hit, then the touched
//                              words are cleared.
//   bitmap_naive.cu:153-181    generateBitmaps (all k bitmaps), then for every
//                              pair a grid-stride AND + __popcll.
//
// MI355X design:
//  * Bitmaps have universe+1 bits (elements reach `universe`; the reference
//    sizes them one bit short, bitmap_naive.cu:68).
//  * bitmap_dynamic SISI is tiled by value: the universe is cut into buckets
//    of 2^bits values (bits chosen so that a bucket holds about 4096
//    elements of the larger set, at most 2^18 values = 32 KiB), and a
//    workgroup takes a run of consecutive buckets; per bucket it clears an
//    LDS bitmap, ORs in the bucket's elements of the SMALLER set (LDS
//    atomics) and probes the bucket's elements of the larger set against it;
//    a thread holds up to 16 elements of each side in registers, and the
//    next bucket's are loaded while the current one is processed.
//    The reference's single global bitmap of the whole universe (512 MiB at
//    2^32, bitmap_dynamic.cu:240-276) becomes a few KiB of LDS; both sets
//    stream once from HBM in order, with no dependent reads.  Bucket bounds
//    in each set are binary searches at the workgroup's start; buckets
//    without elements on either side are skipped.  Counts fold through
//    per-workgroup partials (no reduce kernel).
//  * bitmap_dynamic runs keep the row bitmap in LDS when universe+1 <= 2^20
//    bits (128 KiB of the 160 KiB LDS), else in a per-workgroup global region;
//    each pair's hits are reduced with wave shuffles (one atomic per wave,
//    not one per hit).
//  * bitmap_naive computes all pairs as a tiled AND/popcount "GEMM": a
//    workgroup owns a 64 x 64 block of pairs and streams both operand row
//    blocks through LDS in chunks of 16 words, so every bitmap word loaded is
//    reused 64 times (the reference re-reads both bitmaps for every pair).
#include "si_device.h"
#include "si_internal.h"

namespace {

using namespace si_dev;

// ------------------------------------------------------- bitmap_dynamic --
// One workgroup per run of `per` buckets [v0, v1) of 2^bits values:
// dynamic LDS holds the bucket bitmap (2^(bits-5) words) and the bucket
// bounds of both sets.
template <int NT, int U>
__global__ __launch_bounds__(NT) void k_bm_buckets(const uint32_t* __restrict__ S, uint64_t nS,
                                                   const uint32_t* __restrict__ Q, uint64_t nQ, uint32_t bits,
                                                   uint64_t nb, uint32_t per, uint64_t* __restrict__ partials,
                                                   uint32_t* __restrict__ done, uint64_t* __restrict__ out64,
                                                   uint32_t* __restrict__ out32, uint32_t prio) {
    extern __shared__ uint4 dyn[];
    __shared__ uint64_t red[NT / WAVE];
    uint32_t* bm = reinterpret_cast<uint32_t*>(dyn);
    const uint32_t words = 1u << (bits - 5), low = (1u << bits) - 1u;
    uint64_t* bnd = reinterpret_cast<uint64_t*>(bm + words);  // S bounds [0, nbd), Q bounds [nbd, 2 nbd)
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
        // buckets with elements on both sides; the next one's elements (up to
        // NT * U of each side) are loaded while the current one is processed
        auto work = [&](uint32_t k) {
            while (k + 1 < nbd && (bnd[k] == bnd[k + 1] || bnd[nbd + k] == bnd[nbd + k + 1])) ++k;
            return k;
        };
        // a scalar base (the bucket's first element) and a 32-bit byte
        // offset per load, clamped to the side's last element (a bucket holds
        // at most 2^30 elements, bits <= 30)
        auto fetch = [&](uint32_t k, uint32_t (&xs)[U], uint32_t (&xq)[U]) {
            const uint64_t s0 = bnd[k], s1 = bnd[k + 1], q0 = bnd[nbd + k], q1 = bnd[nbd + k + 1];
            const char* sb = reinterpret_cast<const char*>(S + uniform_u64(s0));
            const char* qb = reinterpret_cast<const char*>(Q + uniform_u64(q0));
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
        const uint32_t wb = bits - 5;  // word index of a value: bits [5, bits)
        uint32_t hits = 0;             // this thread's matches from its registers
        uint32_t xs[U], xq[U], ys[U], yq[U];
        uint32_t k = work(0);
        if (k + 1 < nbd) fetch(k, xs, xq);
        uint32_t step = 0;
        while (k + 1 < nbd) {
            if (prio) set_wave_prio((step++ + blockIdx.x) & 3u);
            const uint32_t kn = work(k + 1);
            if (kn + 1 < nbd) fetch(kn, ys, yq);
            const uint64_t s0 = bnd[k], s1 = bnd[k + 1], q0 = bnd[nbd + k], q1 = bnd[nbd + k + 1];
            for (uint32_t i = threadIdx.x; i < words / 4; i += NT) dyn[i] = make_uint4(0u, 0u, 0u, 0u);
            __syncthreads();
            const uint32_t sn = uniform_u32(uint32_t(s1 - s0 < CAP ? s1 - s0 : CAP));
            const uint32_t qn = uniform_u32(uint32_t(q1 - q0 < CAP ? q1 - q0 : CAP));
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (threadIdx.x + uint32_t(u) * NT < sn)
                    atomicOr(&bm[__builtin_amdgcn_ubfe(xs[u], 5u, wb)], 1u << (xs[u] & 31));
            for (uint64_t i = s0 + CAP + threadIdx.x; i < s1; i += NT) {  // beyond NT * U: a dense bucket
                const uint32_t x = S[i];
                atomicOr(&bm[(x & low) >> 5], 1u << (x & 31));
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < U; ++u)
                hits += (threadIdx.x + uint32_t(u) * NT < qn) ? __builtin_amdgcn_ubfe(bm[__builtin_amdgcn_ubfe(xq[u], 5u, wb)], xq[u], 1u)
                                                               : 0u;
            for (uint64_t i = q0 + CAP + threadIdx.x; i < q1; i += NT) {
                const uint32_t x = Q[i];
                c += (bm[(x & low) >> 5] >> (x & 31)) & 1u;
            }
            __syncthreads();  // probes done before the next bucket clears the bitmap
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

// MISI run: one workgroup per row a; bitmap of WORDS 32-bit words in LDS
// (LDS_WORDS > 0) or in gbm + blockIdx * words (LDS_WORDS == 0).
template <int NT, int LDS_WORDS, bool SELF>
__global__ __launch_bounds__(NT) void k_bm_run(const uint32_t* __restrict__ base, const uint64_t* __restrict__ offsets,
                                               RunArgs r, uint32_t* __restrict__ counts, uint32_t* __restrict__ gbm,
                                               uint64_t words) {
    __shared__ uint32_t lbm[LDS_WORDS > 0 ? LDS_WORDS : 1];
    uint32_t* bm = LDS_WORDS > 0 ? lbm : gbm + uint64_t(blockIdx.x) * words;
    for (uint64_t i = threadIdx.x; i < words; i += NT) bm[i] = 0;
    __threadfence_block();
    __syncthreads();
    for (uint32_t a = r.aStart + blockIdx.x; a < r.aEnd; a += gridDim.x) {
        const uint64_t oA = offsets[a], nA = offsets[a + 1] - oA;
        if (nA == 0) continue;
        for (uint64_t i = threadIdx.x; i < nA; i += NT) {
            const uint32_t x = base[oA + i];
            atomicOr(&bm[x >> 5], 1u << (x & 31));
        }
        __threadfence_block();
        __syncthreads();
        for (uint32_t b = SELF ? a + 1 : r.bStart; b < r.bEnd; ++b) {
            const uint64_t oB = offsets[b], nB = offsets[b + 1] - oB;
            uint32_t c = 0;
            for (uint64_t i = threadIdx.x; i < nB; i += NT) {
                const uint32_t x = base[oB + i];
                const uint64_t w = x >> 5;
                if (w < words) c += (bm[w] >> (x & 31)) & 1u;
            }
            c = wave_sum_u32(c);
            if (lane_id() == 0 && c) atomicAdd(&counts[run_index<SELF>(r, a, b)], c);
        }
        __syncthreads();
        // clear the words set by row a
        for (uint64_t i = threadIdx.x; i < nA; i += NT) bm[base[oA + i] >> 5] = 0;
        __threadfence_block();
        __syncthreads();
    }
}

// --------------------------------------------------------- bitmap_naive --
// Bitmaps of sets [s0, s0 + ns): row (s - s0) has W64 words of 64 bits.
// Bit rows of sets [s0, s0 + ns): row s is w64 64-bit words (2 * w64 32-bit
// words, same bit order) starting at rows + s * w64.  rows must be zero.  A
// workgroup takes whole sets (grid-stride); its waves split a set's chunks.
__global__ void k_bm_build_rows(const uint32_t* __restrict__ base, const uint64_t* __restrict__ offsets, uint32_t s0,
                                uint32_t ns, uint64_t w64, unsigned long long* __restrict__ rows) {
    const uint32_t nw = blockDim.x / WAVE, w = threadIdx.x / WAVE;
    for (uint32_t s = blockIdx.x; s < ns; s += gridDim.x) {
        const uint64_t o = offsets[s0 + s], n = offsets[s0 + s + 1] - o;
        uint32_t* row = reinterpret_cast<uint32_t*>(rows + uint64_t(s) * w64);
        const uint64_t chunks = (n + WAVE - 1) / WAVE;
        for (uint64_t c = w; c < chunks; c += uint64_t(nw) * BUILD_UNROLL) or_bits_chunks(base + o, n, c, nw, row);
    }
}

// The same rows built in LDS when a row fits (w64 even, at most
// LDS_ROW_BYTES): a workgroup clears the row in LDS, ORs its set's bits in
// (plain stores inside a wave's chunk, atomics at chunk edges) and writes the
// whole row, zeros included, with 16-byte stores, so rows need no clearing.
constexpr uint64_t LDS_ROW_BYTES = 64 * 1024;
constexpr int ROWS_NT = 256;

__global__ __launch_bounds__(ROWS_NT) void k_bm_build_rows_lds(const uint32_t* __restrict__ base,
                                                               const uint64_t* __restrict__ offsets, uint32_t s0,
                                                               uint32_t ns, uint64_t w64,
                                                               unsigned long long* __restrict__ rows) {
    extern __shared__ uint4 lrow4[];
    uint32_t* lrow = reinterpret_cast<uint32_t*>(lrow4);
    constexpr uint32_t nw = ROWS_NT / WAVE;
    const uint32_t w = threadIdx.x / WAVE;
    const uint32_t q4 = uint32_t(w64 / 2);  // 16-byte pieces of a row
    for (uint32_t s = blockIdx.x; s < ns; s += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < q4; i += ROWS_NT) lrow4[i] = uint4{0u, 0u, 0u, 0u};
        __syncthreads();
        const uint64_t o = offsets[s0 + s], n = offsets[s0 + s + 1] - o;
        const uint64_t chunks = (n + WAVE - 1) / WAVE;
        if (n < 8 * w64) {
            // fewer than 4 elements per 32-bit word on average: lanes rarely
            // share a word, so one LDS atomic OR per element beats combining
            // the wave's bits first
            for (uint64_t c = w; c < chunks; c += uint64_t(nw) * BUILD_UNROLL) {
                uint32_t x[BUILD_UNROLL];
                bool ok[BUILD_UNROLL];
#pragma unroll
                for (int u = 0; u < BUILD_UNROLL; ++u) {
                    const uint64_t i = (c + uint64_t(u) * nw) * WAVE + lane_id();
                    ok[u] = i < n;
                    x[u] = ok[u] ? base[o + i] : 0u;
                }
#pragma unroll
                for (int u = 0; u < BUILD_UNROLL; ++u)
                    if (ok[u]) atomicOr(&lrow[x[u] >> 5], 1u << (x[u] & 31));
            }
        } else {
            for (uint64_t c = w; c < chunks; c += uint64_t(nw) * BUILD_UNROLL) or_bits_chunks(base + o, n, c, nw, lrow);
        }
        __syncthreads();
        uint4* dst = reinterpret_cast<uint4*>(rows + uint64_t(s) * w64);
        for (uint32_t i = threadIdx.x; i < q4; i += ROWS_NT) dst[i] = lrow4[i];
        __syncthreads();  // the row is read out before the next set clears it
    }
}

// AND + popcount over bit rows: C(i, j) = popcount(R_i & R_j) for i in row
// block [i0, i0+ni) of RA and j in [j0, j0+nj) of RB.
//   TRI: write counts[tri_index(n, i, j)] for j > i only (RA == RB).
//   else dense: C[(i - i0) * ld + (j - j0)].
template <bool TRI>
__global__ __launch_bounds__(256) void k_popcount_pairs(const unsigned long long* __restrict__ RA, uint32_t ni,
                                                        const unsigned long long* __restrict__ RB, uint32_t nj,
                                                        uint64_t w64, uint32_t tilesJ, uint32_t* __restrict__ out,
                                                        uint64_t ld) {
    constexpr int T = 64, KC = 16;
    __shared__ unsigned long long As[T][KC + 1];
    __shared__ unsigned long long Bs[T][KC + 1];
    uint32_t ti, tj;
    if (TRI) {
        tri_decode(blockIdx.x, tilesJ + 1, ti, tj);  // tiles ti <= tj encoded as ti < tj+1
        tj -= 1;
    } else {
        ti = blockIdx.x / tilesJ;
        tj = blockIdx.x % tilesJ;
    }
    const uint32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    uint32_t acc[4][4] = {};
    const uint32_t lr = threadIdx.x >> 2, lc = (threadIdx.x & 3) * 4;  // loader: row, 4 words
    for (uint64_t k0 = 0; k0 < w64; k0 += KC) {
        {
            const uint32_t gi = ti * T + lr, gj = tj * T + lr;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint64_t kk = k0 + lc + q;
                As[lr][lc + q] = (gi < ni && kk < w64) ? RA[uint64_t(gi) * w64 + kk] : 0ull;
                Bs[lr][lc + q] = (gj < nj && kk < w64) ? RB[uint64_t(gj) * w64 + kk] : 0ull;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
            unsigned long long a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[ty + 16 * r][kk];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = Bs[tx + 16 * c][kk];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] += __popcll(a[r] & b[c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t i = ti * T + ty + 16 * r;
        if (i >= ni) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t j = tj * T + tx + 16 * c;
            if (j >= nj) continue;
            if (TRI) {
                if (j > i) out[tri_index(ld, i, j)] = acc[r][c];
            } else {
                out[uint64_t(i) * ld + j] = acc[r][c];
            }
        }
    }
}

constexpr int NT = 256;

}  // namespace

namespace si_impl {

static uint64_t bitmap_words32(si_dataset* ds) { return (uint64_t(std::max(ds->max_elem, ds->universe)) >> 5) + 1; }

int bitmap_pair_count(si_dataset* ds, uint32_t a, uint32_t b, const si_params* p, uint64_t* d_out, uint32_t* d_out32) {
    si_context* ctx = ds->ctx;
    uint32_t s = a, q = b;  // build the smaller set, probe the larger
    if (ds->sizes[s] > ds->sizes[q]) std::swap(s, q);
    const uint64_t nS = ds->sizes[s], nQ = ds->sizes[q];
    if (nS == 0 || nQ == 0) {
        return zero_outputs(ctx, d_out, d_out32);
    }
    // buckets of 2^bits values holding about 4096 elements of the larger set
    // (what the workgroup's registers take at once)
    const uint64_t vmax = uint64_t(std::max(ds->max_elem, ds->universe));
    uint32_t lu = 0, ln = 0;
    while ((uint64_t(1) << lu) <= vmax) ++lu;  // values fit in lu bits
    while ((uint64_t(1) << (ln + 1)) <= nQ) ++ln;
    int bits = int(lu) - int(ln) + 12;
    if (p && p->reserved[1]) bits = int(p->reserved[1]);  // tuning sweeps
    bits = bits < 10 ? 10 : (bits > 18 ? 18 : bits);
    const uint64_t nb = (vmax >> bits) + 1;
    const uint64_t maxGrid = (p && p->blocks) ? p->blocks : uint64_t(ctx->num_cus) * 4;
    uint64_t per = (nb + maxGrid - 1) / maxGrid;
    // the bucket bounds share the workgroup's 64 KiB of dynamic LDS with the
    // bitmap: a small --blocks gets more workgroups than it asked for
    const size_t bmBytes = (size_t(1) << (bits - 5)) * 4;
    const uint64_t perMax = (64 * 1024 - bmBytes) / 16 - 1;
    if (per > perMax) per = perMax;
    const uint64_t g = (nb + per - 1) / per;
    const size_t lds = bmBytes + size_t(2) * (per + 1) * 8;
    void* ws = nullptr;
    SI_TRY(small_workspace(ctx, 256 + g * 8, &ws));
    KernelScope ks(ctx, "bitmap_sisi_buckets");
    hipLaunchKernelGGL((k_bm_buckets<NT, 16>), dim3(uint32_t(g)), dim3(NT), lds, ctx->stream, ds->d_elems + ds->offsets[s],
                       nS, ds->d_elems + ds->offsets[q], nQ, uint32_t(bits), nb, uint32_t(per),
                       reinterpret_cast<uint64_t*>(static_cast<char*>(ws) + 256), static_cast<uint32_t*>(ws), d_out,
                       d_out32, uint32_t(ctx->sisi_prio));
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int bitmap_run(si_dataset* ds, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts) {
    si_context* ctx = ds->ctx;
    const bool self = (A.id == B.id);
    RunArgs r{A.start, A.end, A.id, B.start, B.end, B.id, n};
    const uint64_t words = bitmap_words32(ds);
    uint64_t grid = (p && p->blocks) ? p->blocks : uint64_t(ctx->num_cus) * 4;
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, A.end - A.start));
    KernelScope ks(ctx, "bitmap_run");
    constexpr int SMALL = 8192, LARGE = 32768;  // 32 KiB / 128 KiB of LDS
    if (words <= SMALL) {
        if (self)
            hipLaunchKernelGGL((k_bm_run<NT, SMALL, true>), dim3(uint32_t(grid)), dim3(NT), 0, ctx->stream, ds->d_elems,
                               ds->d_offsets, r, d_counts, nullptr, words);
        else
            hipLaunchKernelGGL((k_bm_run<NT, SMALL, false>), dim3(uint32_t(grid)), dim3(NT), 0, ctx->stream, ds->d_elems,
                               ds->d_offsets, r, d_counts, nullptr, words);
    } else if (words <= LARGE) {
        if (self)
            hipLaunchKernelGGL((k_bm_run<NT, LARGE, true>), dim3(uint32_t(grid)), dim3(NT), 0, ctx->stream, ds->d_elems,
                               ds->d_offsets, r, d_counts, nullptr, words);
        else
            hipLaunchKernelGGL((k_bm_run<NT, LARGE, false>), dim3(uint32_t(grid)), dim3(NT), 0, ctx->stream, ds->d_elems,
                               ds->d_offsets, r, d_counts, nullptr, words);
    } else {
        const uint64_t budget = 4ull << 30;
        grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, budget / (words * 4)));
        SI_TRY(ctx->ws_aux.ensure(grid * words * 4));
        if (self)
            hipLaunchKernelGGL((k_bm_run<NT, 0, true>), dim3(uint32_t(grid)), dim3(NT), 0, ctx->stream, ds->d_elems,
                               ds->d_offsets, r, d_counts, ctx->ws_aux.as<uint32_t>(), words);
        else
            hipLaunchKernelGGL((k_bm_run<NT, 0, false>), dim3(uint32_t(grid)), dim3(NT), 0, ctx->stream, ds->d_elems,
                               ds->d_offsets, r, d_counts, ctx->ws_aux.as<uint32_t>(), words);
    }
    SI_HIP(hipGetLastError());
    return SI_OK;
}

// Bit rows of sets [s0, s0+ns) into `rows` (ns x w64 words): every word is
// written when a row fits in LDS; otherwise the rows are zeroed first (by
// the caller when zero is false) and the set bits ORed in.
bool bit_rows_in_lds(uint64_t w64) { return w64 % 2 == 0 && w64 * 8 <= LDS_ROW_BYTES; }

int build_bit_rows(si_dataset* ds, uint32_t s0, uint32_t ns, uint64_t w64, unsigned long long* rows, bool zero) {
    si_context* ctx = ds->ctx;
    if (ns == 0) return SI_OK;
    if (bit_rows_in_lds(w64)) {
        KernelScope ks(ctx, "bitmap_build_rows");
        const uint32_t grid = std::min<uint32_t>(ns, uint32_t(ctx->num_cus) * 8);
        hipLaunchKernelGGL(k_bm_build_rows_lds, dim3(grid), dim3(ROWS_NT), size_t(w64 * 8), ctx->stream, ds->d_elems,
                           ds->d_offsets, s0, ns, w64, rows);
        SI_HIP(hipGetLastError());
        return SI_OK;
    }
    if (zero) SI_HIP(hipMemsetAsync(rows, 0, uint64_t(ns) * w64 * 8, ctx->stream));
    KernelScope ks(ctx, "bitmap_build_rows");
    const uint32_t grid = std::min<uint32_t>(ns, uint32_t(ctx->num_cus) * 16);
    hipLaunchKernelGGL(k_bm_build_rows, dim3(grid), dim3(NT), 0, ctx->stream, ds->d_elems, ds->d_offsets, s0, ns, w64,
                       rows);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int popcount_pairs(si_context* ctx, const unsigned long long* RA, uint32_t ni, const unsigned long long* RB, uint32_t nj,
                   uint64_t w64, bool tri, uint32_t* out, uint64_t ld) {
    const uint32_t ti = (ni + 63) / 64, tj = (nj + 63) / 64;
    const uint64_t blocks = tri ? uint64_t(tj) * (tj + 1) / 2 : uint64_t(ti) * tj;
    if (blocks == 0) return SI_OK;
    if (blocks > 0x7FFFFFFFull) return fail(SI_ERR_RANGE, "too many pair tiles");
    KernelScope ks(ctx, "popcount_pairs");
    if (tri)
        hipLaunchKernelGGL((k_popcount_pairs<true>), dim3(uint32_t(blocks)), dim3(256), 0, ctx->stream, RA, ni, RB, nj, w64,
                           tj, out, ld);
    else
        hipLaunchKernelGGL((k_popcount_pairs<false>), dim3(uint32_t(blocks)), dim3(256), 0, ctx->stream, RA, ni, RB, nj,
                           w64, tj, out, ld);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

// bitmap_naive: all pairs of the collection, counts[triangular_index(k, a, b)].
int bitmap_all_pairs(si_dataset* ds, const si_params*, uint32_t* d_counts) {
    si_context* ctx = ds->ctx;
    const uint64_t w64 = (uint64_t(std::max(ds->max_elem, ds->universe)) >> 6) + 1;
    const uint64_t bytes = uint64_t(ds->k) * w64 * 8;
    SI_TRY(ctx->ws_aux.ensure(std::max<uint64_t>(bytes, 64)));
    auto* rows = ctx->ws_aux.as<unsigned long long>();
    SI_TRY(build_bit_rows(ds, 0, ds->k, w64, rows));
    return popcount_pairs(ctx, rows, ds->k, rows, ds->k, w64, true, d_counts, ds->k);
}

}  // namespace si_impl
