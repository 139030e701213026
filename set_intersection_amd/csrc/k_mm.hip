// Bitmap-GEMM ("mm") kernels for gfx950: pairwise intersection counts as the
// product of 0/1 incidence matrices on the matrix cores.
//
// Reference behaviour (src/matrix_multiplication.cu): the host fills dense
// float 0/1 matrices of tile A ([A.len][universe+1], populateBinaryMatrix,
// :211-220) and of tile B transposed (:222-231), copies them over PCIe and
// calls cublasSgemm (:161-165); C[i][j] = |A_i ∩ B_j| as fp32, exact while
// counts stay below 2^24.
//
// MI355X design:
//  * Incidence matrices are built on the device from the sorted lists
//    (one byte per bit, rows padded to 128, K padded to 64); nothing crosses
//    PCIe but the counts.
//  * C = M_A · M_Bᵀ with v_mfma_i32_16x16x64_i8: both operands are rows of
//    row-major incidence matrices (K contiguous), so A and B fragments are the
//    same 16-byte row slices.  int32 accumulation is exact below 2^31.
//  * A 256-thread workgroup computes a 128 x 128 tile (2 x 2 waves of 64 x 64,
//    4 x 4 MFMA blocks each); 64-byte K steps are staged through LDS with the
//    next step's global loads issued before the current step's MFMAs.
//  * For a self-join tile (A == B) only tiles on or above the diagonal are
//    computed and each is stored twice (C and Cᵀ), halving the MFMA work.
//  * The FP32 comparison variant uses v_mfma_f32_16x16x32_fp8_fp8 on the same
//    matrices built with the fp8 (e4m3) encoding of 1.0 (0x38): exact only
//    while every count is below 2^24.
#include <type_traits>

#include "si_device.h"
#include "si_internal.h"

namespace {

using namespace si_dev;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int TILE = 128;
constexpr int KS = 64;          // bytes of K per step
constexpr int LDS_STRIDE = 80;  // padded row stride (bytes) in LDS

// rows [s0, s0+ns) of the collection -> mat[(s - s0) * Kp + x] = one
__global__ void k_mm_build(const uint32_t* __restrict__ base, const uint64_t* __restrict__ offsets, uint32_t s0,
                           uint32_t ns, uint64_t Kp, uint8_t one, uint8_t* __restrict__ mat) {
    for (uint32_t s = blockIdx.y; s < ns; s += gridDim.y) {
        const uint64_t o = offsets[s0 + s], n = offsets[s0 + s + 1] - o;
        uint8_t* row = mat + uint64_t(s) * Kp;
        for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
            const uint32_t x = base[o + i];
            if (x < Kp) row[x] = one;
        }
    }
}

template <bool F32>
__global__ __launch_bounds__(256) void k_mm_gemm(const uint8_t* __restrict__ MA, uint32_t ni,
                                                 const uint8_t* __restrict__ MB, uint32_t nj, uint64_t Kp, bool self,
                                                 uint32_t tilesN, uint32_t* __restrict__ C, uint64_t ld) {
    __shared__ __attribute__((aligned(16))) uint8_t As[TILE * LDS_STRIDE];
    __shared__ __attribute__((aligned(16))) uint8_t Bs[TILE * LDS_STRIDE];
    uint32_t tm, tn;
    if (self) {
        tri_decode(blockIdx.x, tilesN + 1, tm, tn);
        tn -= 1;
    } else {
        tm = blockIdx.x / tilesN;
        tn = blockIdx.x % tilesN;
    }
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t wm = w >> 1, wn = w & 1;
    const uint8_t* gA = MA + uint64_t(tm) * TILE * Kp;
    const uint8_t* gB = MB + uint64_t(tn) * TILE * Kp;

    // loader: 2 x 16 B of A and of B per thread per K step
    const uint32_t q0 = tid, q1 = tid + 256u;
    const uint64_t offA0 = uint64_t(q0 >> 2) * Kp + (q0 & 3) * 16, offA1 = uint64_t(q1 >> 2) * Kp + (q1 & 3) * 16;
    const uint32_t lds0 = (q0 >> 2) * LDS_STRIDE + (q0 & 3) * 16, lds1 = (q1 >> 2) * LDS_STRIDE + (q1 & 3) * 16;
    uint4 ra0 = *reinterpret_cast<const uint4*>(gA + offA0), ra1 = *reinterpret_cast<const uint4*>(gA + offA1);
    uint4 rb0 = *reinterpret_cast<const uint4*>(gB + offA0), rb1 = *reinterpret_cast<const uint4*>(gB + offA1);

    using Acc = typename std::conditional<F32, v4f, v4i>::type;
    Acc acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    const uint32_t r16 = lane & 15, kq = lane >> 4;
    for (uint64_t k0 = 0; k0 < Kp; k0 += KS) {
        *reinterpret_cast<uint4*>(&As[lds0]) = ra0;
        *reinterpret_cast<uint4*>(&As[lds1]) = ra1;
        *reinterpret_cast<uint4*>(&Bs[lds0]) = rb0;
        *reinterpret_cast<uint4*>(&Bs[lds1]) = rb1;
        __syncthreads();
        if (k0 + KS < Kp) {
            const uint64_t k1 = k0 + KS;
            ra0 = *reinterpret_cast<const uint4*>(gA + offA0 + k1);
            ra1 = *reinterpret_cast<const uint4*>(gA + offA1 + k1);
            rb0 = *reinterpret_cast<const uint4*>(gB + offA0 + k1);
            rb1 = *reinterpret_cast<const uint4*>(gB + offA1 + k1);
        }
        if constexpr (!F32) {
            v4i a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                a[i] = *reinterpret_cast<const v4i*>(&As[(wm * 64 + i * 16 + r16) * LDS_STRIDE + kq * 16]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                b[j] = *reinterpret_cast<const v4i*>(&Bs[(wn * 64 + j * 16 + r16) * LDS_STRIDE + kq * 16]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], acc[i][j], 0, 0, 0);
        } else {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                long a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    a[i] = *reinterpret_cast<const long*>(&As[(wm * 64 + i * 16 + r16) * LDS_STRIDE + s * 32 + kq * 8]);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    b[j] = *reinterpret_cast<const long*>(&Bs[(wn * 64 + j * 16 + r16) * LDS_STRIDE + s * 32 + kq * 8]);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // lane holds column (lane & 15) and rows 4 * (lane >> 4) + v of each block
    const bool mirror = self && tm != tn;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint32_t row = tm * TILE + wm * 64 + i * 16 + 4 * (lane >> 4) + v;
                const uint32_t col = tn * TILE + wn * 64 + j * 16 + (lane & 15);
                const uint32_t val = uint32_t(acc[i][j][v]);
                if (row < ni && col < nj) {
                    C[uint64_t(row) * ld + col] = val;
                    if (mirror) C[uint64_t(col) * ld + row] = val;
                }
            }
}

}  // namespace

namespace si_impl {

// Dense counts of tile pair (A, B): C[(i - A.start) * ld + (j - B.start)].
int mm_tile_counts(si_dataset* ds, si_tile A, si_tile B, int32_t variant, uint64_t ld, uint32_t* d_C) {
    si_context* ctx = ds->ctx;
    const uint32_t ni = A.end - A.start, nj = B.end - B.start;
    if (ni == 0 || nj == 0) return SI_OK;
    const bool self = (A.start == B.start && A.end == B.end);
    const uint64_t K = uint64_t(std::max(ds->max_elem, ds->universe)) + 1;
    if (variant == SI_MM_POPCOUNT) {
        const uint64_t w64 = (K + 63) / 64;
        const uint64_t rows = self ? ni : uint64_t(ni) + nj;
        SI_TRY(ctx->ws_aux.ensure(rows * w64 * 8));
        auto* RA = ctx->ws_aux.as<unsigned long long>();
        auto* RB = self ? RA : RA + uint64_t(ni) * w64;
        SI_TRY(build_bit_rows(ds, A.start, ni, w64, RA));
        if (!self) SI_TRY(build_bit_rows(ds, B.start, nj, w64, RB));
        return popcount_pairs(ctx, RA, ni, RB, nj, w64, false, d_C, ld);
    }
    const uint64_t Kp = (K + KS - 1) / KS * KS;
    const uint64_t rA = (uint64_t(ni) + TILE - 1) / TILE * TILE, rB = (uint64_t(nj) + TILE - 1) / TILE * TILE;
    const uint64_t bytes = (self ? rA : rA + rB) * Kp;
    SI_TRY(ctx->ws_aux.ensure(bytes));
    uint8_t* MA = ctx->ws_aux.as<uint8_t>();
    uint8_t* MB = self ? MA : MA + rA * Kp;
    const uint8_t one = variant == SI_MM_F32 ? 0x38 : 0x01;  // fp8 e4m3 1.0 / int8 1
    SI_HIP(hipMemsetAsync(MA, 0, bytes, ctx->stream));
    {
        KernelScope ks(ctx, "mm_build");
        hipLaunchKernelGGL(k_mm_build, dim3(4, std::min<uint32_t>(ni, 65535)), dim3(256), 0, ctx->stream, ds->d_elems,
                           ds->d_offsets, A.start, ni, Kp, one, MA);
        if (!self)
            hipLaunchKernelGGL(k_mm_build, dim3(4, std::min<uint32_t>(nj, 65535)), dim3(256), 0, ctx->stream, ds->d_elems,
                               ds->d_offsets, B.start, nj, Kp, one, MB);
        SI_HIP(hipGetLastError());
    }
    const uint32_t tM = uint32_t(rA / TILE), tN = uint32_t(rB / TILE);
    const uint64_t blocks = self ? uint64_t(tN) * (tN + 1) / 2 : uint64_t(tM) * tN;
    if (blocks > 0x7FFFFFFFull) return fail(SI_ERR_RANGE, "too many GEMM tiles");
    KernelScope ks(ctx, variant == SI_MM_F32 ? "mm_gemm_f32" : "mm_gemm_i8");
    if (variant == SI_MM_F32)
        hipLaunchKernelGGL((k_mm_gemm<true>), dim3(uint32_t(blocks)), dim3(256), 0, ctx->stream, MA, ni, MB, nj, Kp, self,
                           tN, d_C, ld);
    else
        hipLaunchKernelGGL((k_mm_gemm<false>), dim3(uint32_t(blocks)), dim3(256), 0, ctx->stream, MA, ni, MB, nj, Kp,
                           self, tN, d_C, ld);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

}  // namespace si_impl
