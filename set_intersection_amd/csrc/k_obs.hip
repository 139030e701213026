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

// SISI: probes P spread over the grid, searched in S.
template <int NT, int NS>
__global__ __launch_bounds__(NT) void k_obs_pair(const uint32_t* __restrict__ base, uint64_t oS, uint64_t nS,
                                                 uint64_t oP, uint64_t nP, uint64_t* __restrict__ partials,
                                                 uint32_t* __restrict__ done, uint64_t* __restrict__ out64,
                                                 uint32_t* __restrict__ out32) {
    __shared__ uint32_t cache[NS];
    __shared__ uint64_t red[NT / WAVE];
    const uint32_t* S = base + oS;
    const uint32_t* P = base + oP;
    load_samples<NT, NS>(cache, S, nS);
    __syncthreads();
    uint64_t c = 0;
    const uint64_t stride = uint64_t(gridDim.x) * NT;
    for (uint64_t i = uint64_t(blockIdx.x) * NT + threadIdx.x; i < nP; i += stride) c += sample_search<NS>(S, nS, cache, P[i]);
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

constexpr int PAIR_NT = 256;
constexpr int PAIR_NS = 4096;

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
    uint64_t grid = (nP + PAIR_NT * 4 - 1) / (PAIR_NT * 4);
    const uint64_t maxGrid = (p && p->blocks) ? p->blocks : uint64_t(ctx->num_cus) * 8;
    grid = std::max<uint64_t>(1, std::min(grid, maxGrid));
    void* ws = nullptr;
    SI_TRY(small_workspace(ctx, 256 + grid * 8, &ws));
    KernelScope ks(ctx, "obs_sisi_count");
    hipLaunchKernelGGL((k_obs_pair<PAIR_NT, PAIR_NS>), dim3(uint32_t(grid)), dim3(PAIR_NT), 0, ctx->stream, ds->d_elems,
                       oS, nS, oP, nP, reinterpret_cast<uint64_t*>(static_cast<char*>(ws) + 256),
                       static_cast<uint32_t*>(ws), d_out, d_out32);
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
