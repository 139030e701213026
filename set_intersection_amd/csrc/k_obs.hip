// Optimized binary search ("obs") kernels for gfx950.
//
// Reference behaviour (src/obs.cu): for a pair (a, b), evenly spaced samples
// A[t*|A|/W] of the searched set are cached in shared memory (obs.cu:296,
// 375); every element of the probed set is first located between two cached
// samples (phase 1, obs.cu:312-328 / 390-407), then binary-searched in global
// memory inside that bracket (phase 2, obs.cu:331-345 / 410-424).  warpBased /
// blockBased differ in whether a warp or a block owns a row a (obs.cu:290,
// 369); --path applies the search inside merge-path partitions (obs.cu:438;
// here k_path.hip).
//
// MI355X design:
//  * Sample positions and bracket bounds use 64-bit products, so sets of any
//    size are exact (obs.cu:375 overflows 32 bits once |A| * blockDim >= 2^32).
//  * SISI spreads the probes of the pair over the whole grid; every
//    workgroup keeps its own LDS copy of NS samples of the searched set, so a
//    probe costs log2(NS) LDS steps + log2(|S|/NS) global steps.  The searched
//    set is the larger one (roles picked by size; counts are symmetric).
//  * MISI runs keep the reference's ownership (a workgroup, or a single wave
//    for --warp, owns a row a and walks b over tile B) and reduce each pair's
//    count with wave shuffles, one atomic per wave instead of per thread.
#include "si_device.h"
#include "si_internal.h"

namespace {

using namespace si_dev;

template <int NT, int NS>
__device__ __forceinline__ void load_samples(uint32_t* cache, const uint32_t* __restrict__ S, uint64_t nS) {
    for (uint32_t t = threadIdx.x; t < NS; t += NT) cache[t] = S[(uint64_t(t) * nS) / NS];
}

// samples[t] = S[t * nS / NS]: gathered once per launch, so the workgroups
// of k_obs_pair and k_obs_mat read 16 KiB of contiguous samples instead of NS
// scattered lines.
template <int NS>
__global__ void k_gather_samples(const uint32_t* __restrict__ S, uint64_t nS, uint32_t* __restrict__ samples) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < NS) samples[t] = S[(uint64_t(t) * nS) / NS];
}

// SISI: probes P spread over the grid, searched in S, R per thread at a
// time (sample_search_multi); the LDS samples are copied from `samples`
// (k_gather_samples), 16 KiB of contiguous lines per workgroup.
template <int NT, int NS, int R>
__global__ __launch_bounds__(NT) void k_obs_pair(const uint32_t* __restrict__ base, uint64_t oS, uint64_t nS,
                                                 uint64_t oP, uint64_t nP, const uint32_t* __restrict__ samples,
                                                 uint64_t* __restrict__ partials, uint32_t* __restrict__ done,
                                                 uint64_t* __restrict__ out64, uint32_t* __restrict__ out32,
                                                 uint32_t prio) {
    __shared__ alignas(16) uint32_t cache[NS];
    __shared__ uint64_t red[NT / WAVE];
    const uint32_t* S = base + oS;
    const uint32_t* P = base + oP;
    for (uint32_t t = threadIdx.x; t < NS / 4; t += NT)
        reinterpret_cast<uint4*>(cache)[t] = reinterpret_cast<const uint4*>(samples)[t];
    __syncthreads();
    uint64_t c = 0;
    const uint64_t stride = uint64_t(gridDim.x) * NT * R;
    uint32_t step = 0;
    for (uint64_t i0 = uint64_t(blockIdx.x) * NT * R + threadIdx.x; i0 < nP; i0 += stride) {
        if (prio) set_wave_prio((step++ + blockIdx.x) & 3u);
        uint32_t x[R];
        bool v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint64_t i = i0 + uint64_t(r) * NT;
            v[r] = i < nP;
            x[r] = v[r] ? P[i] : 0u;
        }
        c += uint32_t(__builtin_popcount(sample_search_multi<NS, R>(S, nS, cache, x, v)));
    }
    if (prio) set_wave_prio(0);
    grid_fold_u64<NT>(c, partials, done, out64, out32, red);
}

// MISI run: a group (NT threads: a workgroup, or one wave for --warp) owns row
// a of tile A and searches the elements of every b of tile B in a.
template <int NT, bool SELF>
__global__ __launch_bounds__(NT) void k_obs_run(const uint32_t* __restrict__ base, const uint64_t* __restrict__ offsets,
                                                RunArgs r, uint32_t* __restrict__ counts) {
    __shared__ uint32_t cache[NT];
    for (uint32_t a = r.aStart + blockIdx.x; a < r.aEnd; a += gridDim.x) {
        const uint64_t oA = offsets[a], nA = offsets[a + 1] - oA;
        if (nA == 0) continue;  // uniform across the group
        const uint32_t* S = base + oA;
        load_samples<NT, NT>(cache, S, nA);
        __syncthreads();
        for (uint32_t b = SELF ? a + 1 : r.bStart; b < r.bEnd; ++b) {
            const uint64_t oB = offsets[b], nB = offsets[b + 1] - oB;
            uint32_t c = 0;
            for (uint64_t i = threadIdx.x; i < nB; i += NT) c += sample_search<NT>(S, nA, cache, base[oB + i]);
            c = wave_sum_u32(c);
            if (lane_id() == 0 && c) atomicAdd(&counts[run_index<SELF>(r, a, b)], c);
        }
        __syncthreads();
    }
}

// SISI materialisation by search, for pairs whose sizes differ widely (the
// merge reads every element of the larger set; this reads about
// log2(|S| / NS) of its lines per probe).  Persistent workgroups take chunks
// of the probed set P by ticket (done[1]): chunk s is NT * R consecutive
// elements, R per thread searched in S together (sample_search_multi); the
// chunk's matches are ranked by ballots and a scan of the (key, wave)
// counts, and land after the matches of chunks 0..s-1 (decoupled look-back
// over the chunks) in P's order, which is ascending.  A workgroup leaves
// after a ticket past the last chunk; the last one to leave (done[2]) clears
// the look-back states and the tickets, when every ticket has been drawn and
// every chunk placed.
template <int NT, int NS, int R>
__global__ __launch_bounds__(NT) void k_obs_mat(const uint32_t* __restrict__ base, uint64_t oS, uint64_t nS,
                                                uint64_t oP, uint64_t nP, const uint32_t* __restrict__ samples,
                                                uint32_t nseg, uint32_t* __restrict__ done,
                                                uint64_t* __restrict__ out, uint32_t* __restrict__ list,
                                                uint64_t* __restrict__ state) {
    constexpr int NW = NT / WAVE;
    __shared__ alignas(16) uint32_t cache[NS];
    __shared__ uint32_t rank[R * NW];  // matches per (key, wave), then their exclusive offsets
    __shared__ uint32_t shTicket, shTotal;
    __shared__ uint64_t shPrefix;
    const uint32_t* S = base + oS;
    const uint32_t* P = base + oP;
    for (uint32_t t = threadIdx.x; t < NS / 4; t += NT)
        reinterpret_cast<uint4*>(cache)[t] = reinterpret_cast<const uint4*>(samples)[t];
    const uint32_t w = threadIdx.x / WAVE, lane = lane_id();
    while (true) {
        if (threadIdx.x == 0) shTicket = atomicAdd(&done[1], 1u);
        __syncthreads();
        const uint32_t s = shTicket;
        if (s >= nseg) break;
        // key r of this thread: element c0 + r * NT + thread (coalesced loads)
        const uint64_t c0 = uint64_t(s) * (NT * R);
        uint32_t x[R];
        bool v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint64_t i = c0 + uint64_t(r) * NT + threadIdx.x;
            v[r] = i < nP;
            x[r] = v[r] ? P[i] : 0u;
        }
        const uint32_t hit = sample_search_multi<NS, R>(S, nS, cache, x, v);
        uint64_t m[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            m[r] = __ballot((hit >> r) & 1u);
            if (lane == 0) rank[r * NW + w] = uint32_t(__popcll(m[r]));
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (int k = 0; k < R * NW; ++k) {
                const uint32_t c = rank[k];
                rank[k] = run;
                run += c;
            }
            shTotal = run;
        }
        __syncthreads();
        if (threadIdx.x < WAVE) {
            const uint64_t pre = lookback_exclusive(state, s, shTotal);
            if (lane == 0) {
                shPrefix = pre;
                if (s == nseg - 1) *out = pre + shTotal;
            }
        }
        __syncthreads();
        const uint64_t pre = shPrefix;
        for (int r = 0; r < R; ++r) {  // (fully unrolled by the compiler: R is a constant)
            if ((m[r] >> lane) & 1ull) {
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(uint32_t(m[r] >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo(uint32_t(m[r]), 0u));
                list[pre + rank[r * NW + w] + below] = x[r];
            }
        }
    }
    // (shTicket may still be read by threads leaving the loop: a variable of its own)
    __shared__ uint32_t shExit;
    if (threadIdx.x == 0) shExit = atomicAdd(&done[2], 1u);
    __syncthreads();
    if (shExit == gridDim.x - 1) {
        for (uint32_t i = threadIdx.x; i < nseg; i += NT)
            __hip_atomic_store(&state[i], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (threadIdx.x == 0) {
            __hip_atomic_store(&done[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&done[2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

constexpr int PAIR_NT = 256;
constexpr int PAIR_NS = 4096;
constexpr int PAIR_R = 4;  // keys per thread of the SISI search (count)
constexpr int MAT_R = 4;   // keys per thread of the search materialisation

}  // namespace

namespace si_impl {

int obs_pair_count(si_dataset* ds, uint32_t a, uint32_t b, const si_params* p, uint64_t* d_out, uint32_t* d_out32) {
    if (p && p->path) return path_pair_count(ds, a, b, SI_TECH_OBS, p, d_out, d_out32);
    si_context* ctx = ds->ctx;
    // search the elements of the smaller set in the larger one
    uint32_t s = a, q = b;
    if (ds->sizes[q] > ds->sizes[s]) std::swap(s, q);
    const uint64_t oS = ds->offsets[s], nS = ds->sizes[s], oP = ds->offsets[q], nP = ds->sizes[q];
    if (nS == 0 || nP == 0) {
        return zero_outputs(ctx, d_out, d_out32);
    }
    uint64_t grid = (nP + PAIR_NT * PAIR_R - 1) / (PAIR_NT * PAIR_R);
    const uint64_t maxGrid = (p && p->blocks) ? p->blocks : uint64_t(ctx->num_cus) * 8;
    grid = std::max<uint64_t>(1, std::min(grid, maxGrid));
    // ticket at byte 0, per-workgroup partials from byte 256, then the samples
    const uint64_t sampleOff = (256 + grid * 8 + 15) & ~uint64_t(15);
    void* ws = nullptr;
    SI_TRY(small_workspace(ctx, sampleOff + PAIR_NS * 4, &ws));
    uint32_t* samples = reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + sampleOff);
    KernelScope ks(ctx, "obs_sisi_count");
    hipLaunchKernelGGL((k_gather_samples<PAIR_NS>), dim3(PAIR_NS / 256), dim3(256), 0, ctx->stream, ds->d_elems + oS,
                       nS, samples);
    hipLaunchKernelGGL((k_obs_pair<PAIR_NT, PAIR_NS, PAIR_R>), dim3(uint32_t(grid)), dim3(PAIR_NT), 0, ctx->stream,
                       ds->d_elems, oS, nS, oP, nP, samples, reinterpret_cast<uint64_t*>(static_cast<char*>(ws) + 256),
                       static_cast<uint32_t*>(ws), d_out, d_out32, uint32_t(ctx->sisi_prio));
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int obs_pair_materialize(si_dataset* ds, uint32_t a, uint32_t b, uint32_t* d_out, uint64_t* d_count) {
    si_context* ctx = ds->ctx;
    uint32_t s = a, q = b;
    if (ds->sizes[q] > ds->sizes[s]) std::swap(s, q);
    const uint64_t oS = ds->offsets[s], nS = ds->sizes[s], oP = ds->offsets[q], nP = ds->sizes[q];
    if (nS == 0 || nP == 0) return zero_outputs(ctx, d_count, nullptr);
    const uint64_t chunks = (nP + PAIR_NT * MAT_R - 1) / (PAIR_NT * MAT_R);
    if (chunks > 0xFFFFFFFFull) return fail(SI_ERR_RANGE, "search materialisation: too many chunks");
    const uint32_t nseg = uint32_t(chunks);
    // tickets at bytes 0-11, the gathered samples from byte 256
    void* ws = nullptr;
    SI_TRY(small_workspace(ctx, 256 + PAIR_NS * 4, &ws));
    uint32_t* samples = reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + 256);
    // look-back states: zeroed once when (re)allocated, left zeroed by every launch
    if (ctx->ws_lookback.bytes < uint64_t(nseg) * 8) {
        SI_TRY(ctx->ws_lookback.ensure(std::max<uint64_t>(uint64_t(nseg) * 8, 1 << 14)));
        SI_HIP(hipMemsetAsync(ctx->ws_lookback.ptr, 0, ctx->ws_lookback.bytes, ctx->stream));
    }
    const uint32_t grid = uint32_t(std::min<uint64_t>(nseg, uint64_t(ctx->num_cus) * 8));
    KernelScope ks(ctx, "obs_sisi_materialize");
    hipLaunchKernelGGL((k_gather_samples<PAIR_NS>), dim3(PAIR_NS / 256), dim3(256), 0, ctx->stream, ds->d_elems + oS,
                       nS, samples);
    hipLaunchKernelGGL((k_obs_mat<PAIR_NT, PAIR_NS, MAT_R>), dim3(grid), dim3(PAIR_NT), 0, ctx->stream, ds->d_elems,
                       oS, nS, oP, nP, samples, nseg, static_cast<uint32_t*>(ws), d_count, d_out,
                       ctx->ws_lookback.as<uint64_t>());
    SI_HIP(hipGetLastError());
    return SI_OK;
}

int obs_run(si_dataset* ds, const si_params* p, si_tile A, si_tile B, uint32_t n, uint32_t* d_counts) {
    if (p && p->path) return path_run(ds, SI_TECH_OBS, p, A, B, n, d_counts);
    si_context* ctx = ds->ctx;
    const bool self = (A.id == B.id);
    RunArgs r{A.start, A.end, A.id, B.start, B.end, B.id, n};
    const bool warp = p && p->warp;
    uint64_t grid = (p && p->blocks) ? p->blocks : uint64_t(ctx->num_cus) * (warp ? 32 : 8);
    grid = std::max<uint64_t>(1, std::min<uint64_t>(grid, A.end - A.start));
    KernelScope ks(ctx, warp ? "obs_run_warp" : "obs_run_block");
#define LAUNCH(NT, S) \
    hipLaunchKernelGGL((k_obs_run<NT, S>), dim3(uint32_t(grid)), dim3(NT), 0, ctx->stream, ds->d_elems, ds->d_offsets, r, d_counts)
    if (warp) {
        if (self) LAUNCH(64, true); else LAUNCH(64, false);
    } else {
        if (self) LAUNCH(256, true); else LAUNCH(256, false);
    }
#undef LAUNCH
    SI_HIP(hipGetLastError());
    return SI_OK;
}

}  // namespace si_impl
