// qg_gram.hip — Gram products C = A^T B for the Rayleigh-Ritz step, K = grid size (10^5-10^7),
// M, N = band counts (<= a few thousand), on the CDNA4 matrix cores.
//
// Inputs are FP32 band-major blocks (column m of A is A + m*lda, K contiguous values); they are
// rounded to BF16 in registers and multiplied with v_mfma_f32_32x32x16_bf16 (FP32 accumulate);
// the split variant carries x = hi + lo (two BF16 terms) through three MFMAs per step.
// Each wavefront owns a 64 x 64 output tile (2 x 2 MFMA blocks, 64 accumulators per lane) and
// one K-chunk (split-K); a lane loads 8 consecutive K values of one row of A / column of B per
// MFMA (32 B of FP32, fully contiguous along K).  Partial tiles go to an FP32 workspace and a
// second kernel sums the K-chunks in FP64 into the FP64 result (column stride ldc).
#include <algorithm>
#include "qg_internal.hpp"

namespace qg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGramTile = 64;
constexpr int kGramKStep = 16;

__device__ __forceinline__ bf16x8 to_bf16x8(f32x4 a, f32x4 b) {
    bf16x8 r;
    r[0] = (__bf16)a[0]; r[1] = (__bf16)a[1]; r[2] = (__bf16)a[2]; r[3] = (__bf16)a[3];
    r[4] = (__bf16)b[0]; r[5] = (__bf16)b[1]; r[6] = (__bf16)b[2]; r[7] = (__bf16)b[3];
    return r;
}

// split x = hi + lo with hi = bf16(x), lo = bf16(x - hi): products of (hi + lo) keep ~16
// mantissa bits, hi*hi + hi*lo + lo*hi drops only the lo*lo term (2^-16 relative)
__device__ __forceinline__ void split_bf16x8(f32x4 a, f32x4 b, bf16x8& hi, bf16x8& lo) {
    hi = to_bf16x8(a, b);
    f32x4 ra, rb;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ra[j] = a[j] - (float)hi[j];
        rb[j] = b[j] - (float)hi[j + 4];
    }
    lo = to_bf16x8(ra, rb);
}

template <bool SPLIT>
__device__ __forceinline__ f32x16 mfma_acc(const bf16x8& ah, const bf16x8& al, const bf16x8& bh, const bf16x8& bl,
                                           f32x16 c) {
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, c, 0, 0, 0);
    if constexpr (SPLIT) {
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, c, 0, 0, 0);
    }
    return c;
}

// one wavefront per (64 x 64 tile, K chunk); partial[chunk][n][m] (column-major tiles)
template <bool SPLIT>
__global__ __launch_bounds__(64) void gram_bf16_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                       int64_t K, int M, int N, int64_t lda, int64_t ldb,
                                                       int64_t kchunk, float* __restrict__ partial) {
    const int tiles_m = (M + kGramTile - 1) / kGramTile;
    const int tiles_n = (N + kGramTile - 1) / kGramTile;
    const int t = blockIdx.x % (tiles_m * tiles_n);
    const int chunk = blockIdx.x / (tiles_m * tiles_n);
    const int tm = t % tiles_m, tn = t / tiles_m;
    const int lane = threadIdx.x;
    const int r32 = lane & 31, kh = lane >> 5;
    const int64_t k0 = (int64_t)chunk * kchunk;
    const int64_t k1 = min(K, k0 + kchunk);
    // rows of A and columns of B this lane feeds (two 32-blocks each)
    const int m0 = tm * kGramTile + r32, m1 = m0 + 32;
    const int n0 = tn * kGramTile + r32, n1 = n0 + 32;
    const float* a0 = A + (int64_t)min(m0, M - 1) * lda;
    const float* a1 = A + (int64_t)min(m1, M - 1) * lda;
    const float* b0 = B + (int64_t)min(n0, N - 1) * ldb;
    const float* b1 = B + (int64_t)min(n1, N - 1) * ldb;
    const bool va0 = m0 < M, va1 = m1 < M, vb0 = n0 < N, vb1 = n1 < N;
    f32x16 c00 = {}, c01 = {}, c10 = {}, c11 = {};
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
    // main loop over full K steps (16 values per MFMA, 8 per lane)
    int64_t k = k0;
    const bool aligned = ((lda | ldb) & 3) == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0;
    if (aligned) {
        for (; k + kGramKStep <= k1 && (k & 3) == 0; k += kGramKStep) {
            const int64_t kk = k + kh * 8;
            const f32x4* pa0 = reinterpret_cast<const f32x4*>(a0 + kk);
            const f32x4* pa1 = reinterpret_cast<const f32x4*>(a1 + kk);
            const f32x4* pb0 = reinterpret_cast<const f32x4*>(b0 + kk);
            const f32x4* pb1 = reinterpret_cast<const f32x4*>(b1 + kk);
            bf16x8 ha0, la0, ha1, la1, hb0, lb0, hb1, lb1;
            split_bf16x8(va0 ? pa0[0] : z4, va0 ? pa0[1] : z4, ha0, la0);
            split_bf16x8(va1 ? pa1[0] : z4, va1 ? pa1[1] : z4, ha1, la1);
            split_bf16x8(vb0 ? pb0[0] : z4, vb0 ? pb0[1] : z4, hb0, lb0);
            split_bf16x8(vb1 ? pb1[0] : z4, vb1 ? pb1[1] : z4, hb1, lb1);
            c00 = mfma_acc<SPLIT>(ha0, la0, hb0, lb0, c00);
            c01 = mfma_acc<SPLIT>(ha0, la0, hb1, lb1, c01);
            c10 = mfma_acc<SPLIT>(ha1, la1, hb0, lb0, c10);
            c11 = mfma_acc<SPLIT>(ha1, la1, hb1, lb1, c11);
        }
    }
    // remainder (and unaligned inputs): element-wise loads with K bounds
    for (; k < k1; k += kGramKStep) {
        float ta0[8], ta1[8], tb0[8], tb1[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t kk = k + kh * 8 + j;
            const bool ok = kk < k1;
            ta0[j] = (ok && va0) ? a0[kk] : 0.f;
            ta1[j] = (ok && va1) ? a1[kk] : 0.f;
            tb0[j] = (ok && vb0) ? b0[kk] : 0.f;
            tb1[j] = (ok && vb1) ? b1[kk] : 0.f;
        }
        const f32x4 x0 = {ta0[0], ta0[1], ta0[2], ta0[3]}, x1 = {ta0[4], ta0[5], ta0[6], ta0[7]};
        const f32x4 y0 = {ta1[0], ta1[1], ta1[2], ta1[3]}, y1 = {ta1[4], ta1[5], ta1[6], ta1[7]};
        const f32x4 u0 = {tb0[0], tb0[1], tb0[2], tb0[3]}, u1 = {tb0[4], tb0[5], tb0[6], tb0[7]};
        const f32x4 w0 = {tb1[0], tb1[1], tb1[2], tb1[3]}, w1 = {tb1[4], tb1[5], tb1[6], tb1[7]};
        bf16x8 ha0, la0, ha1, la1, hb0, lb0, hb1, lb1;
        split_bf16x8(x0, x1, ha0, la0);
        split_bf16x8(y0, y1, ha1, la1);
        split_bf16x8(u0, u1, hb0, lb0);
        split_bf16x8(w0, w1, hb1, lb1);
        c00 = mfma_acc<SPLIT>(ha0, la0, hb0, lb0, c00);
        c01 = mfma_acc<SPLIT>(ha0, la0, hb1, lb1, c01);
        c10 = mfma_acc<SPLIT>(ha1, la1, hb0, lb0, c10);
        c11 = mfma_acc<SPLIT>(ha1, la1, hb1, lb1, c11);
    }
    // 32x32 accumulator layout: value r of lane l is (row 8*(r/4) + 4*(l/32) + r%4, column l%32)
    float* P = partial + (int64_t)chunk * M * N;
    auto put = [&](const f32x16& c, int mb, int nb) {
        const int n = tn * kGramTile + nb + r32;
        if (n >= N) return;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int m = tm * kGramTile + mb + 8 * (r / 4) + 4 * kh + (r % 4);
            if (m < M) P[(int64_t)n * M + m] = c[r];
        }
    };
    put(c00, 0, 0);
    put(c01, 0, 32);
    put(c10, 32, 0);
    put(c11, 32, 32);
}

__global__ void gram_reduce_kernel(const float* __restrict__ partial, int nchunk, int M, int N, double* C, int ldc) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)M * N) return;
    const int m = (int)(i % M), n = (int)(i / M);
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += (double)partial[(int64_t)c * M * N + i];
    C[m + (int64_t)ldc * n] = s;
}

static int64_t gram_chunks(int64_t K, int M, int N) {
    const int64_t tiles = (int64_t)((M + kGramTile - 1) / kGramTile) * ((N + kGramTile - 1) / kGramTile);
    // ~4096 wavefronts, chunks of at least 4096 K values
    int64_t nch = std::max<int64_t>(1, 4096 / std::max<int64_t>(1, tiles));
    nch = std::min<int64_t>(nch, std::max<int64_t>(1, K / 4096));
    return nch;
}

size_t gram_bf16_workspace(int64_t K, int M, int N) { return (size_t)gram_chunks(K, M, N) * M * N; }

void gram_bf16(const float* A, const float* B, int64_t K, int M, int N, int64_t lda, int64_t ldb, double* C,
               int ldc, float* partial, hipStream_t s, bool split) {
    const int64_t nch = gram_chunks(K, M, N);
    int64_t kchunk = (K + nch - 1) / nch;
    kchunk = (kchunk + kGramKStep - 1) / kGramKStep * kGramKStep;
    const int64_t tiles = (int64_t)((M + kGramTile - 1) / kGramTile) * ((N + kGramTile - 1) / kGramTile);
    if (split)
        hipLaunchKernelGGL(gram_bf16_kernel<true>, dim3((unsigned)(tiles * nch)), dim3(64), 0, s, A, B, K, M, N, lda,
                           ldb, kchunk, partial);
    else
        hipLaunchKernelGGL(gram_bf16_kernel<false>, dim3((unsigned)(tiles * nch)), dim3(64), 0, s, A, B, K, M, N, lda,
                           ldb, kchunk, partial);
    QG_HIP(hipGetLastError());
    const int64_t mn = (int64_t)M * N;
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, partial, (int)nch, M,
                       N, C, ldc);
    QG_HIP(hipGetLastError());
}

}  // namespace qg

using namespace qg;

extern "C" int qg_gram_bf16(const float* A, const float* B, int64_t K, int M, int N, int64_t lda, int64_t ldb,
                            double* C, int ldc, int split, void* stream) {
    return guarded([&] {
        QG_REQUIRE(A && B && C && K > 0 && M > 0 && N > 0 && lda >= K && ldb >= K && ldc >= M, "bad arguments");
        DevBuf<float> ws;
        ws.alloc(gram_bf16_workspace(K, M, N));
        gram_bf16(A, B, K, M, N, lda, ldb, C, ldc, ws.p, as_stream(stream), split != 0);
        QG_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}
