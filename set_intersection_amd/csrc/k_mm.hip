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
//  * Operands stay bit-packed: one 64-bit word per 64 universe values per set
//    (8 KiB per set at 2^16), built on the device from the sorted lists; the
//    bytes the matrix cores consume are made in registers right before each
//    MFMA, so the operand traffic is an eighth of a byte matrix's.
//  * C = M_A · M_Bᵀ with v_mfma_i32_16x16x64_i8: a lane's 16 operand bytes
//    are 16 universe values of one row, taken from a 32-bit word of its
//    bitmap with one shift and one mask per 4 bytes (byte i of dword j is
//    bit 8i + j of the shifted word); A and B use the same value-to-K
//    assignment, so the products pair up.  int32 accumulation is exact
//    below 2^31.  The K step is software-pipelined: the next fragment is
//    expanded while the current fragment's MFMAs issue.
//  * A two-wave workgroup computes a 128 x 128 tile (each wave 128 x 64,
//    8 x 4 MFMA blocks: an expanded fragment feeds 4 or 8 MFMAs).  Chunks of
//    8 words (512 universe values) of the 256 operand rows are staged through
//    double-buffered LDS (odd dword row stride: the 16 rows a fragment reads
//    hit distinct banks); the next chunk's global loads are issued before the
//    current chunk's MFMAs.
//  * Whole tiles fill complete rounds of the resident workgroup slots; the
//    tiles of a last partial round are split along K into up to 16 parts
//    added atomically into zeroed C, so no round runs mostly empty.
//  * For a self-join tile (A == B) only tiles on or above the diagonal are
//    computed and each is stored twice (C and Cᵀ), halving the MFMA work.
//  * The FP32-accumulate comparison variant expands the same bits to the fp8
//    (e4m3) encoding of 1.0 (0x38) and uses v_mfma_f32_16x16x32_fp8_fp8:
//    exact only while every count is below 2^24.
//  * The fp4 variant feeds v_mfma_scale_f32_16x16x128_f8f6f4 with e2m1
//    nibbles (scales 2^0): 128 universe values per MFMA, twice the K of the
//    int8 instruction in the same issue slots.  A lane's 32 nibbles come from
//    one 32-bit word with one or two operations per dword: nibble i of dword
//    j holds bit 4i + j, masked in place where the e2m1 code of the single set
//    bit is a power of two (bit 0: 0.5, bit 1: 1.0, bit 2: 2.0; bit 3 is the
//    sign, so it is shifted to bit 1), and B takes the reciprocal weights
//    (2.0, 1.0, 0.5, 1.0), so every product of two set bits is exactly 1.0.
//    fp32 accumulation: exact while every count is below 2^24.
#include <type_traits>

#include "si_device.h"
#include "si_internal.h"

namespace {

using namespace si_dev;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int TILE = 128;
constexpr int KC = 8;              // 64-bit words per staged chunk
constexpr int RS = 2 * KC + 1;     // LDS row stride in dwords (odd)

// A lane's 16 operand bytes: dword j (j = 0..3) of lane half p (p = 0, 1)
// holds bits 8i + 4p + j (i = 0..3) of a 32-bit bitmap word in its bytes i,
// as 0/1 (int8) or 0/0x38 (fp8 1.0).  Any fixed assignment of universe
// values to MFMA K positions is correct as long as A and B use the same one;
// this one costs a shift and a mask per dword.
// fp4 operands (see above): A's dword j carries weight 2^(j-1) (j < 3) or 1,
// B's the reciprocal
__device__ __forceinline__ v4i expand_a4(uint32_t x) {
    return v4i{int(x & 0x11111111u), int(x & 0x22222222u), int(x & 0x44444444u), int((x >> 2) & 0x22222222u)};
}
__device__ __forceinline__ v4i expand_b4(uint32_t x) {
    const uint32_t y = x >> 2;
    return v4i{int((x << 2) & 0x44444444u), int(x & 0x22222222u), int(y & 0x11111111u), int(y & 0x22222222u)};
}
typedef int v8i __attribute__((ext_vector_type(8)));
__device__ __forceinline__ v8i f4_operand(const v4i& f) { return v8i{f[0], f[1], f[2], f[3], 0, 0, 0, 0}; }

enum { MODE_I8 = 0, MODE_F8 = 1, MODE_F4 = 2, MODE_I8W = 3 };

// Weighted int8 operands (MODE_I8W): the same bits as expand32, but A's dword
// j carries a set bit as the byte value 2^j (a mask in place, no shift) and
// B's as 2^(3-j) (bit j moved to bit 3-j), so the product of two set bits is
// 8 in every dword and the count is the accumulated sum / 8.  A fragment is
// expanded twice as often as a B fragment (8 against 4 per K step), so the
// shift-free side is A.  Exact while every count is below 2^28.
__device__ __forceinline__ v4i expand_w_a(uint32_t x, uint32_t p) {
    const uint32_t y = x >> (4 * p);
    return v4i{int(y & 0x01010101u), int(y & 0x02020202u), int(y & 0x04040404u), int(y & 0x08080808u)};
}
__device__ __forceinline__ v4i expand_w_b(uint32_t x, uint32_t p) {
    const uint32_t y = x >> (4 * p);
    return v4i{int((y << 3) & 0x08080808u), int((y << 1) & 0x04040404u), int((y >> 1) & 0x02020202u),
               int((y >> 3) & 0x01010101u)};
}

template <bool F8>
__device__ __forceinline__ v4i expand32(uint32_t x, uint32_t p) {
    const uint32_t y = x >> (4 * p);
    v4i r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t v = (y >> j) & 0x01010101u;
        if (F8) v *= 0x38u;
        r[j] = int(v);
    }
    return r;
}

// Work unit u of a GEMM launch: units [0, nfull) are whole tiles; each of
// the ntail tiles after them is split into nsplit K ranges (units nfull + ...)
// whose partial counts are added atomically into zeroed C.
struct MmGrid {
    uint32_t nfull, ntail, nsplit, tilesN;
    bool self;
};

__device__ __forceinline__ void mm_tile_of(const MmGrid& g, uint32_t t, uint32_t& tm, uint32_t& tn) {
    if (g.self) {
        tri_decode(t, g.tilesN + 1, tm, tn);
        tn -= 1;
    } else {
        tm = t / g.tilesN;
        tn = t % g.tilesN;
    }
}

// C(i, j) = |row i ∩ row j| for a 128 x 128 tile of bit rows RA (ni rows) x
// RB (nj rows), w64 words per row (a multiple of KC; rows padded to a
// multiple of 128 with zeros).  Two waves: wave w computes all 128 rows x
// columns [64w, 64w + 64) (8 x 4 MFMA blocks), so an expanded fragment feeds
// 4 (A) or 8 (B) MFMAs.  self: RA == RB, tiles on or above the diagonal only,
// each stored to C and Cᵀ.
template <int MODE>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mm_bits(const uint64_t* __restrict__ RA, uint32_t ni,
                                                 const uint64_t* __restrict__ RB, uint32_t nj, uint64_t w64,
                                                 MmGrid g, uint32_t* __restrict__ C, uint64_t ld) {
    __shared__ uint32_t sA[2][TILE * RS];
    __shared__ uint32_t sB[2][TILE * RS];
    const uint32_t u = blockIdx.x;
    uint32_t tm, tn, c0 = 0, c1 = uint32_t(w64 / KC);
    bool split = false;
    if (u < g.nfull) {
        mm_tile_of(g, u, tm, tn);
    } else {
        const uint32_t v = u - g.nfull, part = v % g.nsplit;
        mm_tile_of(g, g.nfull + v / g.nsplit, tm, tn);
        const uint32_t nch = c1;
        c0 = uint32_t(uint64_t(nch) * part / g.nsplit);
        c1 = uint32_t(uint64_t(nch) * (part + 1) / g.nsplit);
        split = true;
    }
    const uint32_t tid = threadIdx.x, lane = tid & 63, wn = tid >> 6;
    const uint64_t* gA = RA + uint64_t(tm) * TILE * w64;
    const uint64_t* gB = RB + uint64_t(tn) * TILE * w64;

    // loader: a chunk is 128 rows x 8 words per side = 512 x 16 B; thread t
    // moves pieces t + 128 p (row = piece / 4, words 2 * (piece % 4) + {0,1})
    uint64_t goff[4];
    uint32_t loff[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t piece = tid + 128u * p, row = piece >> 2, c2 = (piece & 3) * 2;
        goff[p] = uint64_t(row) * w64 + c2;
        loff[p] = row * RS + 2 * c2;
    }
    uint4 ra[4], rb[4];
    auto load = [&](uint64_t k) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            ra[p] = *reinterpret_cast<const uint4*>(gA + goff[p] + k);
            rb[p] = *reinterpret_cast<const uint4*>(gB + goff[p] + k);
        }
    };
    auto store = [&](uint32_t buf) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            uint32_t* da = &sA[buf][loff[p]];
            uint32_t* db = &sB[buf][loff[p]];
            da[0] = ra[p].x; da[1] = ra[p].y; da[2] = ra[p].z; da[3] = ra[p].w;
            db[0] = rb[p].x; db[1] = rb[p].y; db[2] = rb[p].z; db[3] = rb[p].w;
        }
    };

    constexpr bool INT = MODE == MODE_I8 || MODE == MODE_I8W;
    using Acc = typename std::conditional<INT, v4i, v4f>::type;
    Acc acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    // a lane's operand rows: 16 * i + (lane & 15); lane quarter q takes half
    // q & 1 of the bits of dword q >> 1 of each 64-bit word
    const uint32_t r16 = lane & 15, q = lane >> 4, p = q & 1, h = q >> 1;
    const uint32_t rowA = r16 * RS + h, rowB = (wn * 64 + r16) * RS + h;
    load(uint64_t(c0) * KC);
    store(0);
    __syncthreads();
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t buf = (c - c0) & 1;
        if (c + 1 < c1) load(uint64_t(c + 1) * KC);
        const uint32_t* dA = sA[buf];
        const uint32_t* dB = sB[buf];
        if constexpr (MODE == MODE_F4) {
            // as the int8 pipeline below, with K steps of 128 universe values
            // (4 dwords of a row; lane quarter q takes dword q)
            const uint32_t* qA = dA + (r16 * RS + q);
            const uint32_t* qB = dB + ((wn * 64 + r16) * RS + q);
            uint32_t wa[8], wb[4];
#pragma unroll
            for (int i = 0; i < 8; ++i) wa[i] = qA[i * 16 * RS];
#pragma unroll
            for (int j = 0; j < 4; ++j) wb[j] = qB[j * 16 * RS];
            v4i fb[4], fbn[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) fb[j] = expand_b4(wb[j]);
            v4i fan = expand_a4(wa[0]);
#pragma unroll 2
            for (int kk = 0; kk < KC / 2; ++kk) {
                const bool more = kk + 1 < KC / 2;
                const int k4 = more ? 4 * (kk + 1) : 4 * kk;
                uint32_t na[8], nb[4];
#pragma unroll
                for (int i = 0; i < 8; ++i) na[i] = qA[i * 16 * RS + k4];
#pragma unroll
                for (int j = 0; j < 4; ++j) nb[j] = qB[j * 16 * RS + k4];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const v4i fa = fan;
                    if (i < 7) fan = expand_a4(wa[i + 1]);
                    else fan = expand_a4(na[0]);
                    if (i >= 3 && i < 7) fbn[i - 3] = expand_b4(nb[i - 3]);
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                            f4_operand(fa), f4_operand(fb[j]), acc[i][j], 4, 4, 0, 127, 0, 127);
                    // (no explicit interleaving: the compiler's own schedule
                    // measured 8% faster than MFMA/VALU group barriers here)
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) wa[i] = na[i];
#pragma unroll
                for (int j = 0; j < 4; ++j) fb[j] = fbn[j];
            }
        } else if constexpr (INT) {
            // software pipeline over the 8 A fragments of a K step: while the
            // 4 MFMAs of fragment i issue, fragment i+1 (and from i = 3 on one
            // B fragment of the next K step) is expanded, so the expansion
            // runs in the shadow of the matrix pipe
            uint32_t wa[8], wb[4];
#pragma unroll
            for (int i = 0; i < 8; ++i) wa[i] = dA[rowA + i * 16 * RS];
#pragma unroll
            for (int j = 0; j < 4; ++j) wb[j] = dB[rowB + j * 16 * RS];
            v4i fb[4], fbn[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) fb[j] = MODE == MODE_I8W ? expand_w_b(wb[j], p) : expand32<false>(wb[j], p);
            v4i fan = MODE == MODE_I8W ? expand_w_a(wa[0], p) : expand32<false>(wa[0], p);
#pragma unroll 2
            for (int kk = 0; kk < KC; ++kk) {
                const bool more = kk + 1 < KC;
                uint32_t na[8], nb[4];
                const int k2 = more ? 2 * (kk + 1) : 2 * kk;
#pragma unroll
                for (int i = 0; i < 8; ++i) na[i] = dA[rowA + i * 16 * RS + k2];
#pragma unroll
                for (int j = 0; j < 4; ++j) nb[j] = dB[rowB + j * 16 * RS + k2];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const v4i fa = fan;
                    const uint32_t wn1 = i < 7 ? wa[i + 1] : na[0];
                    fan = MODE == MODE_I8W ? expand_w_a(wn1, p) : expand32<false>(wn1, p);
                    if (i >= 3 && i < 7)
                        fbn[i - 3] = MODE == MODE_I8W ? expand_w_b(nb[i - 3], p) : expand32<false>(nb[i - 3], p);
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa, fb[j], acc[i][j], 0, 0, 0);
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) wa[i] = na[i];
#pragma unroll
                for (int j = 0; j < 4; ++j) fb[j] = fbn[j];
            }
        } else {
            for (int kk = 0; kk < KC; ++kk) {
                v4i fb[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) fb[j] = expand32<true>(dB[rowB + j * 16 * RS + 2 * kk], p);
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const v4i fa = expand32<true>(dA[rowA + i * 16 * RS + 2 * kk], p);
                    // the fp8 MFMA takes 8 bytes per lane: the fragment's two
                    // halves are two K sub-steps (the same for A and B)
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const long av = (long(uint32_t(fa[2 * hh + 1])) << 32) | uint32_t(fa[2 * hh]);
                            const long bv = (long(uint32_t(fb[j][2 * hh + 1])) << 32) | uint32_t(fb[j][2 * hh]);
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(av, bv, acc[i][j], 0, 0, 0);
                        }
                }
            }
        }
        if (c + 1 < c1) store(buf ^ 1);
        __syncthreads();
    }

    // lane holds column (lane & 15) and rows 4 * (lane >> 4) + v of each block
    const bool mirror = g.self && tm != tn;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint32_t row = tm * TILE + i * 16 + 4 * (lane >> 4) + v;
                const uint32_t col = tn * TILE + wn * 64 + j * 16 + (lane & 15);
                const uint32_t val = MODE == MODE_I8W ? uint32_t(acc[i][j][v]) >> 3 : uint32_t(acc[i][j][v]);
                if (row < ni && col < nj) {
                    if (split) {
                        atomicAdd(&C[uint64_t(row) * ld + col], val);
                        if (mirror) atomicAdd(&C[uint64_t(col) * ld + row], val);
                    } else {
                        C[uint64_t(row) * ld + col] = val;
                        if (mirror) C[uint64_t(col) * ld + row] = val;
                    }
                }
            }
}

// Zeroes the C regions of the tail tiles [t0, t0 + n) (and their mirrors).
__global__ void k_mm_zero_tail(MmGrid g, uint32_t t0, uint32_t ni, uint32_t nj, uint32_t* __restrict__ C,
                               uint64_t ld) {
    uint32_t tm, tn;
    mm_tile_of(g, t0 + blockIdx.x, tm, tn);
    for (uint32_t e = threadIdx.x; e < TILE * TILE; e += blockDim.x) {
        const uint32_t row = tm * TILE + e / TILE, col = tn * TILE + e % TILE;
        if (row < ni && col < nj) {
            C[uint64_t(row) * ld + col] = 0;
            if (g.self && tm != tn) C[uint64_t(col) * ld + row] = 0;
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
    // fp32 accumulation is exact only while every count (at most K) is below 2^24
    if ((variant == SI_MM_F32 || variant == SI_MM_FP4) && K > (uint64_t(1) << 24))
        return fail(SI_ERR_RANGE, "universe above 2^24: FP32-accumulated counts could round (use SI_MM_I8)");
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
    // bit rows: w64 a multiple of the staged chunk, rows padded to the tile
    const uint64_t w64 = (K + 64 * KC - 1) / (64 * KC) * KC;
    const uint64_t rA = (uint64_t(ni) + TILE - 1) / TILE * TILE, rB = (uint64_t(nj) + TILE - 1) / TILE * TILE;
    const uint64_t rows = self ? rA : rA + rB;
    SI_TRY(ctx->ws_aux.ensure(rows * w64 * 8));
    auto* RA = ctx->ws_aux.as<unsigned long long>();
    auto* RB = self ? RA : RA + rA * w64;
    if (bit_rows_in_lds(w64)) {
        // the builder writes whole rows: clear only the padding rows
        if (rA > ni) SI_HIP(hipMemsetAsync(RA + uint64_t(ni) * w64, 0, (rA - ni) * w64 * 8, ctx->stream));
        if (!self && rB > nj) SI_HIP(hipMemsetAsync(RB + uint64_t(nj) * w64, 0, (rB - nj) * w64 * 8, ctx->stream));
    } else {
        SI_HIP(hipMemsetAsync(RA, 0, rows * w64 * 8, ctx->stream));
    }
    SI_TRY(build_bit_rows(ds, A.start, ni, w64, RA, false));
    if (!self) SI_TRY(build_bit_rows(ds, B.start, nj, w64, RB, false));
    const uint32_t tM = uint32_t(rA / TILE), tN = uint32_t(rB / TILE);
    const uint64_t tiles = self ? uint64_t(tN) * (tN + 1) / 2 : uint64_t(tM) * tN;
    if (tiles > 0x7FFFFFFFull) return fail(SI_ERR_RANGE, "too many GEMM tiles");
    // whole tiles fill complete rounds of the resident slots (4 two-wave
    // workgroups per CU); the tiles of a last partial round are split along K
    const uint32_t slots = uint32_t(ctx->num_cus) * 4;
    MmGrid g{uint32_t(tiles), 0, 1, tN, self};
    const uint32_t nch = uint32_t(w64 / KC);
    if (tiles > slots && tiles % slots) {
        const uint32_t tail = uint32_t(tiles % slots);
        uint32_t sp = 1;
        while (sp * 2 * tail <= slots && sp * 2 <= 16 && sp * 2 <= nch) sp *= 2;
        if (sp > 1) {
            g.nfull = uint32_t(tiles) - tail;
            g.ntail = tail;
            g.nsplit = sp;
            hipLaunchKernelGGL(k_mm_zero_tail, dim3(tail), dim3(256), 0, ctx->stream, g, g.nfull, ni, nj, d_C, ld);
        }
    }
    const uint32_t units = g.nfull + g.ntail * g.nsplit;
    const uint64_t* bA = reinterpret_cast<const uint64_t*>(RA);
    const uint64_t* bB = reinterpret_cast<const uint64_t*>(RB);
    KernelScope ks(ctx, variant == SI_MM_F32 ? "mm_gemm_f32" : variant == SI_MM_FP4 ? "mm_gemm_fp4" : "mm_gemm_i8");
    if (variant == SI_MM_F32)
        hipLaunchKernelGGL((k_mm_bits<MODE_F8>), dim3(units), dim3(128), 0, ctx->stream, bA, ni, bB, nj, w64, g, d_C, ld);
    else if (variant == SI_MM_FP4)
        hipLaunchKernelGGL((k_mm_bits<MODE_F4>), dim3(units), dim3(128), 0, ctx->stream, bA, ni, bB, nj, w64, g, d_C, ld);
    else if (ds->max_size < (1u << 28))  // weighted operands: 8 x count below 2^31
        hipLaunchKernelGGL((k_mm_bits<MODE_I8W>), dim3(units), dim3(128), 0, ctx->stream, bA, ni, bB, nj, w64, g, d_C, ld);
    else
        hipLaunchKernelGGL((k_mm_bits<MODE_I8>), dim3(units), dim3(128), 0, ctx->stream, bA, ni, bB, nj, w64, g, d_C, ld);
    SI_HIP(hipGetLastError());
    return SI_OK;
}

}  // namespace si_impl
