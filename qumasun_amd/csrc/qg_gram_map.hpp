// qg_gram_map.hpp — tiling of the split-BF16 Gram kernel (qg_gram.hip), shared with its host
// check (tests/harness/gram_bounds.cpp): tile sizes, the K chunking, which global values a
// thread stages where, and which output element an accumulator value is.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace qg {

constexpr int kGT = 256;                  // tile rows and columns
constexpr int kGK = 32;                   // K values per stage
constexpr int kGRow = kGK + 8;            // BF16 per LDS row
constexpr int kGThreads = 512;
constexpr int kGLoads = kGT * kGK / 4 / kGThreads;   // float4 loads per thread per operand (4)

// Staging map shared by the kernel and its host check: float4 load i of thread tid reads row
// `row` of the tile (A row tm*kGT + row, B column tn*kGT + row) at K offset k4 within the stage
// and lands at LDS element row * kGRow + k4 of the hi / lo tiles.
__host__ __device__ __forceinline__ void gram_stage_map(int tid, int i, int& row, int& k4) {
    const int q = tid + kGThreads * i;
    row = q >> 3;
    k4 = (q & 7) * 4;
}
// K range [k0, k1) of chunk c for chunk length kchunk (a multiple of kGK)
__host__ __device__ __forceinline__ void gram_chunk_range(int64_t c, int64_t kchunk, int64_t K, int64_t& k0,
                                                          int64_t& k1) {
    k0 = c * kchunk;
    k1 = k0 + kchunk < K ? k0 + kchunk : K;
}
// tile, chunk of workgroup b
__host__ __device__ __forceinline__ void gram_block_map(int b, int M, int N, int& tm, int& tn, int& chunk) {
    const int tiles_m = (M + kGT - 1) / kGT;
    const int tiles_n = (N + kGT - 1) / kGT;
    const int t = b % (tiles_m * tiles_n);
    chunk = b / (tiles_m * tiles_n);
    tm = t % tiles_m;
    tn = t / tiles_m;
}
// output element (m, n) of accumulator value r of block (a, b) of lane `lane` of wavefront `wave`
__host__ __device__ __forceinline__ void gram_out_map(int wave, int lane, int a, int b, int r, int tm, int tn, int& m,
                                                      int& n) {
    const int wm = wave & 3, wn = wave >> 2;
    n = tn * kGT + wn * 128 + b * 32 + (lane & 31);
    m = tm * kGT + wm * 64 + a * 32 + 8 * (r / 4) + (lane >> 5) * 4 + (r % 4);
}


// K chunks of one product: ~512 workgroups (two per CU), chunks of at least 4096 K values, and
// no more partial tiles than the workspace holds (cap floats; 0: unbounded)
inline int64_t gram_bf16_chunks(int64_t K, int M, int N, size_t cap) {
    const int64_t tiles = (int64_t)((M + kGT - 1) / kGT) * ((N + kGT - 1) / kGT);
    int64_t nch = std::max<int64_t>(1, 512 / std::max<int64_t>(1, tiles));
    nch = std::min<int64_t>(nch, std::max<int64_t>(1, K / 4096));
    if (cap > 0) nch = std::min<int64_t>(nch, std::max<int64_t>(1, (int64_t)(cap / ((size_t)M * N))));
    return nch;
}

// K values per chunk for nch chunks: ceil(K / nch) rounded up to whole stages
inline int64_t gram_bf16_kchunk(int64_t K, int64_t nch) {
    const int64_t kc = (K + nch - 1) / nch;
    return (kc + kGK - 1) / kGK * kGK;
}

}  // namespace qg
