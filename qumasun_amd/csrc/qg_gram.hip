// qg_gram.hip — Gram products C = A^T B for the Rayleigh-Ritz step, K = grid size (10^5-10^7),
// M, N = band counts (<= a few thousand), on the CDNA4 matrix cores.
//
// Inputs are FP32 band-major blocks (column m of A is A + m*lda, K contiguous values); they are
// rounded to BF16 when staged in LDS and multiplied with v_mfma_f32_32x32x16_bf16 (FP32
// accumulate); the split variant carries x = hi + lo (two BF16 terms) through three MFMAs per
// step, keeping ~16 mantissa bits.  Split-K: each workgroup owns a 256 x 256 tile of C and one K
// chunk, writes its FP32 partial tile to a workspace, and a second kernel sums the chunks in
// FP64 into the FP64 result (column stride ldc).
#include <algorithm>
#include "qg_internal.hpp"
#include "qg_gram_map.hpp"

namespace qg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));


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

// ------------------------------------------------------------------ tiled kernel ----
// One workgroup of 512 threads (8 wavefronts) per (256 x 256 tile of C, K chunk).  Per stage of
// 32 K values the workgroup stages the tile's 256 rows of A and 256 columns of B in LDS as BF16
// hi (and lo) parts, rows of 32 values padded to 40 (80 bytes: the 16-byte operand reads of one
// wavefront spread over the banks); each wavefront multiplies its 64 x 128 part (2 x 4 blocks of
// 32 x 32, 128 accumulators per lane) with v_mfma_f32_32x32x16_bf16.  The next stage's FP32
// values are loaded into registers while the current one is multiplied (8 consecutive threads
// read one row's 128 contiguous bytes).  Traffic: A is read N/256 times, B M/256 times.
__device__ __forceinline__ void split4(f32x4 x, uint2& hi, uint2& lo) {
    __bf16 h[4], l[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        h[j] = (__bf16)x[j];
        l[j] = (__bf16)(x[j] - (float)h[j]);
    }
    hi = __builtin_bit_cast(uint2, h);
    lo = __builtin_bit_cast(uint2, l);
}

template <bool SPLIT>
__global__ __launch_bounds__(kGThreads) void gram_tile_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                              int64_t K, int M, int N, int64_t lda, int64_t ldb,
                                                              int64_t kchunk, float* __restrict__ partial) {
    extern __shared__ __align__(16) unsigned char gsm[];
    __bf16* sAh = reinterpret_cast<__bf16*>(gsm);
    __bf16* sBh = sAh + kGT * kGRow;
    __bf16* sAl = sBh + kGT * kGRow;
    __bf16* sBl = sAl + kGT * kGRow;
    int tm, tn, chunk;
    gram_block_map(blockIdx.x, M, N, tm, tn, chunk);
    const int tid = threadIdx.x;
    int64_t k0, k1;
    gram_chunk_range(chunk, kchunk, K, k0, k1);
    const bool aligned = ((lda | ldb) & 3) == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0;

    // staging: float4 q = tid + 512 i -> row q / 8, K offset 4 (q % 8)
    const float* pa[kGLoads];
    const float* pb[kGLoads];
    bool va[kGLoads], vb[kGLoads];
    int lo_off[kGLoads];
    int kq[kGLoads];
#pragma unroll
    for (int i = 0; i < kGLoads; ++i) {
        int row, k4;
        gram_stage_map(tid, i, row, k4);
        const int m = tm * kGT + row, n = tn * kGT + row;
        va[i] = m < M;
        vb[i] = n < N;
        pa[i] = A + (int64_t)min(m, M - 1) * lda;
        pb[i] = B + (int64_t)min(n, N - 1) * ldb;
        kq[i] = k4;
        lo_off[i] = row * kGRow + k4;
    }
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
    f32x4 ra[kGLoads], rb[kGLoads];
    auto fetch = [&](const float* p, bool valid, int64_t k) -> f32x4 {
        // 4 values at K index k .. k + 3 of one row (zero past the chunk or the matrix)
        if (!valid) return z4;
        if (aligned && k + 3 < k1) return *reinterpret_cast<const f32x4*>(p + k);
        f32x4 v;
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = (k + j < k1) ? p[k + j] : 0.f;
        return v;
    };
    auto load = [&](int64_t k) {
#pragma unroll
        for (int i = 0; i < kGLoads; ++i) {
            ra[i] = fetch(pa[i], va[i], k + kq[i]);
            rb[i] = fetch(pb[i], vb[i], k + kq[i]);
        }
    };

    // wavefront part: rows 64 wm .. + 64 (2 blocks), columns 128 wn .. + 128 (4 blocks)
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 3, wn = wave >> 2;
    const int r32 = lane & 31, kh = (lane >> 5) * 8;
    f32x16 c[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) c[a][b] = f32x16{};

    load(k0);
    for (int64_t k = k0; k < k1; k += kGK) {
        __syncthreads();   // the previous stage's readers are done
#pragma unroll
        for (int i = 0; i < kGLoads; ++i) {
            uint2 h, l;
            split4(ra[i], h, l);
            *reinterpret_cast<uint2*>(sAh + lo_off[i]) = h;
            if constexpr (SPLIT) *reinterpret_cast<uint2*>(sAl + lo_off[i]) = l;
            split4(rb[i], h, l);
            *reinterpret_cast<uint2*>(sBh + lo_off[i]) = h;
            if constexpr (SPLIT) *reinterpret_cast<uint2*>(sBl + lo_off[i]) = l;
        }
        __syncthreads();
        if (k + kGK < k1) load(k + kGK);
#pragma unroll
        for (int ks = 0; ks < kGK; ks += 16) {
            bf16x8 ah[2], al[2], bh[4], bl[4];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int off = (wm * 64 + a * 32 + r32) * kGRow + ks + kh;
                ah[a] = *reinterpret_cast<const bf16x8*>(sAh + off);
                if constexpr (SPLIT) al[a] = *reinterpret_cast<const bf16x8*>(sAl + off);
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int off = (wn * 128 + b * 32 + r32) * kGRow + ks + kh;
                bh[b] = *reinterpret_cast<const bf16x8*>(sBh + off);
                if constexpr (SPLIT) bl[b] = *reinterpret_cast<const bf16x8*>(sBl + off);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) c[a][b] = mfma_acc<SPLIT>(ah[a], al[a], bh[b], bl[b], c[a][b]);
        }
    }
    // 32 x 32 accumulator layout: value r of lane l is (row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32)
    float* P = partial + (int64_t)chunk * M * N;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                int m, n;
                gram_out_map(wave, lane, a, b, r, tm, tn, m, n);
                if (m < M && n < N) P[(int64_t)n * M + m] = c[a][b][r];
            }
}

__global__ void gram_reduce_kernel(const float* __restrict__ partial, int nchunk, int M, int N, double* C, int ldc) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)M * N) return;
    const int m = (int)(i % M), n = (int)(i / M);
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += (double)partial[(int64_t)c * M * N + i];
    C[m + (int64_t)ldc * n] = s;
}

// Workspace (floats) of one product at the chunk count it would choose unbounded; a caller
// that runs several shapes on one workspace sizes it for one of them, and the others fit their
// chunk count to it.
size_t gram_bf16_workspace(int64_t K, int M, int N) { return (size_t)gram_bf16_chunks(K, M, N, 0) * M * N; }

void gram_bf16(const float* A, const float* B, int64_t K, int M, int N, int64_t lda, int64_t ldb, double* C,
               int ldc, float* partial, size_t partial_floats, hipStream_t s, bool split) {
    const int64_t nch = gram_bf16_chunks(K, M, N, partial_floats);
    QG_REQUIRE((size_t)nch * M * N <= partial_floats, "BF16 Gram workspace too small for one partial tile set");
    const int64_t kchunk = gram_bf16_kchunk(K, nch);
    const int64_t nblk = (int64_t)((M + kGT - 1) / kGT) * ((N + kGT - 1) / kGT) * nch;
    QG_REQUIRE(nblk < (1LL << 31), "too many workgroups");
    const size_t lds = 4 * (size_t)kGT * kGRow * sizeof(__bf16);
    if (split)
        hipLaunchKernelGGL(gram_tile_kernel<true>, dim3((unsigned)nblk), dim3(kGThreads), lds, s, A, B, K, M, N, lda,
                           ldb, kchunk, partial);
    else
        hipLaunchKernelGGL(gram_tile_kernel<false>, dim3((unsigned)nblk), dim3(kGThreads), lds, s, A, B, K, M, N, lda,
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
        gram_bf16(A, B, K, M, N, lda, ldb, C, ldc, ws.p, ws.n, as_stream(stream), split != 0);
        QG_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}
