// Host check of the split-BF16 Gram kernel's tiling (qumasun_amd/csrc/qg_gram_map.hpp, the
// maps the kernel itself uses) for one product shape K x M x N (leading dimensions lda, ldb)
// run on a workspace of cap floats (0: the product's own size):
//   - every global value a thread stages is inside its operand (row < rows, K index < K) and
//     inside its chunk; every LDS slot written is inside the tile;
//   - coverage (when check_cover): per (tile, chunk) every (row, K index) of the chunk is
//     staged exactly once for A and for B;
//   - every partial-tile write is inside the workspace, and each (chunk, m, n) is written once.
// Prints one summary line.
#include "../../qumasun_amd/csrc/qg_gram_map.hpp"
#include <cstdio>
#include <cstdlib>
#include <vector>
using namespace qg;

int main(int argc, char** argv) {
    const int64_t K = atoll(argv[1]);
    const int M = atoi(argv[2]), N = atoi(argv[3]);
    const int64_t lda = atoll(argv[4]), ldb = atoll(argv[5]);
    const size_t cap_in = argc > 6 ? (size_t)atoll(argv[6]) : 0;
    const bool check_cover = argc > 7 ? atoi(argv[7]) != 0 : true;
    const int64_t nch = gram_bf16_chunks(K, M, N, cap_in);
    const size_t cap = cap_in ? cap_in : (size_t)nch * M * N;
    const int64_t kchunk = gram_bf16_kchunk(K, nch);
    const int tiles_m = (M + kGT - 1) / kGT, tiles_n = (N + kGT - 1) / kGT;
    const int64_t nblk = (int64_t)tiles_m * tiles_n * nch;
    int64_t bad_read = 0, bad_lds = 0, bad_cover = 0, bad_write = 0, dup_write = 0, reads = 0;
    std::vector<unsigned char> wrote(check_cover ? (size_t)nch * M * N : 0, 0);
    std::vector<unsigned char> cov;
    for (int64_t b = 0; b < nblk; ++b) {
        int tm, tn, chunk;
        gram_block_map((int)b, M, N, tm, tn, chunk);
        int64_t k0, k1;
        gram_chunk_range(chunk, kchunk, K, k0, k1);
        if (k1 > K || k0 < 0) ++bad_read;
        const int64_t klen = k1 > k0 ? k1 - k0 : 0;
        if (check_cover) cov.assign((size_t)2 * kGT * (klen > 0 ? klen : 1), 0);
        // stages: all of them for coverage, else the first and the last (where k1 cuts)
        std::vector<int64_t> stages;
        for (int64_t k = k0; k < k1; k += kGK)
            if (check_cover || k == k0 || k + kGK >= k1) stages.push_back(k);
        for (int tid = 0; tid < kGThreads; ++tid)
            for (int i = 0; i < kGLoads; ++i) {
                int row, k4;
                gram_stage_map(tid, i, row, k4);
                if (row < 0 || row >= kGT || row * kGRow + k4 + 3 >= kGT * kGRow) ++bad_lds;
                for (int op = 0; op < 2; ++op) {
                    const int rows = op == 0 ? M : N;
                    const int64_t ld = op == 0 ? lda : ldb;
                    const int r = (op == 0 ? tm : tn) * kGT + row;
                    if (r >= rows) continue;   // staged as zeros, nothing read
                    for (int64_t k : stages)
                        for (int j = 0; j < 4; ++j) {
                            const int64_t kk = k + k4 + j;
                            if (!(kk < k1)) continue;   // zero past the chunk
                            ++reads;
                            if (kk < k0 || kk >= K || kk >= ld || (int64_t)r * ld + kk >= (int64_t)rows * ld) ++bad_read;
                            if (check_cover) {
                                unsigned char& c = cov[((size_t)op * kGT + row) * klen + (kk - k0)];
                                if (c) ++bad_cover;
                                c = 1;
                            }
                        }
                }
            }
        if (check_cover)
            for (int op = 0; op < 2; ++op) {
                const int rows = op == 0 ? M : N;
                for (int row = 0; row < kGT; ++row) {
                    if ((op == 0 ? tm : tn) * kGT + row >= rows) continue;
                    for (int64_t kk = 0; kk < klen; ++kk)
                        if (!cov[((size_t)op * kGT + row) * klen + kk]) ++bad_cover;
                }
            }
        for (int wave = 0; wave < kGThreads / 64; ++wave)
            for (int lane = 0; lane < 64; ++lane)
                for (int a = 0; a < 2; ++a)
                    for (int bb = 0; bb < 4; ++bb)
                        for (int r = 0; r < 16; ++r) {
                            int m, n;
                            gram_out_map(wave, lane, a, bb, r, tm, tn, m, n);
                            if (m >= M || n >= N) continue;
                            const size_t idx = (size_t)chunk * M * N + (size_t)n * M + m;
                            if (idx >= cap) ++bad_write;
                            else if (check_cover) {
                                if (wrote[idx]) ++dup_write;
                                wrote[idx] = 1;
                            }
                        }
    }
    int64_t unwritten = 0;
    if (check_cover)
        for (unsigned char w : wrote) unwritten += w == 0;
    printf("chunks %lld kchunk %lld blocks %lld reads %lld bad_read %lld bad_lds %lld bad_cover %lld bad_write %lld "
           "dup_write %lld unwritten %lld\n",
           (long long)nch, (long long)kchunk, (long long)nblk, (long long)reads, (long long)bad_read,
           (long long)bad_lds, (long long)bad_cover, (long long)bad_write, (long long)dup_write, (long long)unwritten);
    return 0;
}
