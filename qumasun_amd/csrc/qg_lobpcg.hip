// qg_lobpcg.hip — device-resident block LOBPCG for the lowest eigenpairs of H on MI355X.
//
// The iteration is the reference's LOBPCG::Eigen_d_multi_mpi (lobpcg_d_multi_mpi.h:884-1572)
// on one GPU: X is scaled by sqrt(dVol) on entry; the basis is [X R P]; the Rayleigh-Ritz
// problem is assembled in the 3N x 3N column-major Sa (H) / Sb (S) layout with upper blocks
// only; the first call of an SCF run uses the 2N basis [X R]; a full Gram refresh happens on
// iteration 0, every refresh_period iterations and after a rollback, otherwise only XtHR,
// RtHR, RtHP and RtR are computed and the other blocks follow the small-matrix recurrences
// of NextMatrix_d (:265-460); the next vectors follow NextVectors2 (:550-579); a norm drift
// |1-||x'||^2| > rollback_limit rolls back (:1307-1346).  Extensions (off by default, so the
// default run follows the reference): a residual tolerance stop, and a preconditioner K
// applied to R (then X^T R and R^T P are computed as GEMMs, since the recurrence identity
// R = AX - XE no longer holds for K R).
//
// Arithmetic: the vector blocks (X, R, P, AX, AR, AP, T) are FP64 in QG_PREC_F64 and FP32
// otherwise; every N x N matrix, the Cholesky factorisation and the generalized eigensolve
// (rocSOLVER dsygvd on the device) are FP64.  Tall-skinny products are plain library GEMMs
// (rocBLAS) except the QG_PREC_MIXED H-matrix Gram products, which use this library's
// split-BF16 MFMA kernel with FP32 accumulation (the overlap matrix stays FP32).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <cmath>
#include <cstdio>
#include <vector>
#include "qg_internal.hpp"

namespace qg {

void hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                       int64_t ld_out, hipStream_t s);
void precond_apply(qg_ham* h, int kind, double shift, int dtype, int nbands, void* r, int64_t ld,
                   hipStream_t s);
void cnn_precond_apply(qg_cnn* c, int dtype, int nbands, void* r, int64_t ld, hipStream_t s);

struct BatchReq {
    qg_ham* h;
    int nbands;
    const void* psi;
    void* out;
};
void batch_hamiltonian_apply(const std::vector<BatchReq>& reqs, int dtype, hipStream_t s,
                             DevBuf<unsigned char>& scratch, DevBuf<double>& coefs);

// Rendezvous of the cells of a batched solve: each cell's solver thread posts its H psi
// request and waits; the last thread to arrive runs every pending request as one batched
// (persistent-kernel) application on its own stream, then wakes the others.  Cells whose
// solve has finished leave the rendezvous.
struct BatchService {
    std::mutex mu;
    std::condition_variable cv;
    int active = 0;
    uint64_t gen = 0;
    struct Req {
        BatchReq r;
        int dtype;
    };
    std::vector<Req> pending;
    std::string error;   // failure of the last generation only (cleared when the next one runs)
    DevBuf<unsigned char> scratch;
    DevBuf<double> coefs;
    int launches = 0;

    // runs with mu held; every active thread is waiting, so pending is stable.  A waiting
    // thread wakes before any later generation can run (that needs its own next request),
    // so the error it reads is its own generation's.
    void execute(hipStream_t s) {
        error.clear();
        try {
            for (int dt : {QG_F32, QG_F64}) {
                std::vector<BatchReq> v;
                for (auto& q : pending)
                    if (q.dtype == dt) v.push_back(q.r);
                if (!v.empty()) {
                    batch_hamiltonian_apply(v, dt, s, scratch, coefs);
                    ++launches;
                }
            }
        } catch (const Error& e) {
            error = e.msg;
        }
        pending.clear();
        ++gen;
        cv.notify_all();
    }
};

static void batch_service_apply(BatchService* b, qg_ham* h, int dtype, int nb, const void* in, void* out,
                                hipStream_t s) {
    QG_HIP(hipStreamSynchronize(s));   // the request's input is complete on this cell's stream
    std::unique_lock<std::mutex> lk(b->mu);
    b->pending.push_back(BatchService::Req{BatchReq{h, nb, in, out}, dtype});
    const uint64_t my = b->gen;
    if ((int)b->pending.size() == b->active) b->execute(s);
    else b->cv.wait(lk, [&] { return b->gen != my; });
    if (!b->error.empty()) throw Error{QG_ERR_HIP, "batched H psi failed: " + b->error};
}

static void batch_service_leave(BatchService* b, hipStream_t s) {
    std::unique_lock<std::mutex> lk(b->mu);
    --b->active;
    if (b->active > 0 && (int)b->pending.size() == b->active) b->execute(s);
}
void gram_bf16(const float* A, const float* B, int64_t K, int M, int N, int64_t lda, int64_t ldb,
               double* C, int ldc, float* partial, size_t partial_floats, hipStream_t s, bool split);
size_t gram_bf16_workspace(int64_t K, int M, int N);

#define QG_BLAS(call)                                                                        \
    do {                                                                                     \
        rocblas_status st_ = (call);                                                         \
        if (st_ != rocblas_status_success)                                                   \
            throw ::qg::Error{QG_ERR_LIB, std::string(#call) + ": " + rocblas_status_to_string(st_)}; \
    } while (0)

// ------------------------------------------------------------------ small kernels ----
template <class T>
__global__ void scale_kernel(T* x, int64_t n, double c) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (T)(x[i] * c);
}

template <class Ts, class Td>
__global__ void convert_kernel(const Ts* a, Td* b, int64_t n, double c) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (Td)((double)a[i] * c);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; ++i) t += red[i];
    return t;
}

// out[k] = sum_i a[k*ld + i] * b[k*ld + i]  (one workgroup per band, FP64 accumulation)
template <class T>
__global__ __launch_bounds__(256) void coldot_kernel(const T* a, const T* b, int64_t ld, int64_t n, double* out) {
    __shared__ double red[4];
    const int k = blockIdx.x;
    const T* pa = a + (int64_t)k * ld;
    const T* pb = b + (int64_t)k * ld;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += (double)pa[i] * (double)pb[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[k] = s;
}

// r[k] = ax[k] - e[k] x[k]; norm2[k] = ||r[k]||^2
template <class T>
__global__ __launch_bounds__(256) void residual_kernel(const T* ax, const T* x, const double* e, T* r, int64_t ld,
                                                       int64_t n, double* norm2) {
    __shared__ double red[4];
    const int k = blockIdx.x;
    const double ek = e[k];
    const T* pa = ax + (int64_t)k * ld;
    const T* px = x + (int64_t)k * ld;
    T* pr = r + (int64_t)k * ld;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const T v = (T)(pa[i] - (T)ek * px[i]);
        pr[i] = v;
        s += (double)v * (double)v;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) norm2[k] = s;
}

// C[m + ldc n] = sum_c W[c][m + N n] (split-K partial Gram products)
template <class T>
__global__ void split_reduce_kernel(const T* W, int nch, int M, int N, double* C, int ldc) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nn = (int64_t)M * N;
    if (i >= nn) return;
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += (double)W[(int64_t)c * nn + i];
    C[(i % M) + (int64_t)ldc * (i / M)] = s;
}

// x_k *= 1 / sqrt(n2[k]) for the columns with n2[k] > 0 (also applied to a partner block y)
template <class T>
__global__ void colscale_kernel(T* x, T* y, const double* n2, int64_t ld, int64_t n) {
    const int k = blockIdx.y;
    const double v = n2[k];
    if (!(v > 0.0)) return;
    const double c = 1.0 / sqrt(v);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[(int64_t)k * ld + i] = (T)(x[(int64_t)k * ld + i] * c);
        if (y) y[(int64_t)k * ld + i] = (T)(y[(int64_t)k * ld + i] * c);
    }
}

// C[:, j] (column stride ldc, FP64) = G[:, j] (N x N, FP32)
__global__ void widen_kernel(const float* G, double* C, int N, int ldc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * N) return;
    C[(i % N) + (int64_t)ldc * (i / N)] = (double)G[i];
}

// XtR = tXtHX - tXtX E (columns scaled by e_j), RtP = tXtHP - E tXtP (rows by e_i)
__global__ void direct_xr_rp_kernel(int N, double* XtR, double* RtP, int ld, const double* tXtHX,
                                    const double* tXtX, const double* tXtHP, const double* tXtP,
                                    const double* e) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= N) return;
    XtR[i + (int64_t)ld * j] = tXtHX[i + (int64_t)N * j] - tXtX[i + (int64_t)N * j] * e[j];
    RtP[i + (int64_t)ld * j] = tXtHP[i + (int64_t)N * j] - e[i] * tXtP[i + (int64_t)N * j];
}

static int grid_for(int64_t n) { return (int)std::min<int64_t>(4096, (n + 255) / 256); }

// K chunks of the split-K Gram (each chunk summed by the library in the block precision, the
// chunks in FP64).  FP32: chunks of at most ~6,750 terms (256 at 120^3) keep the Gram rounding
// low enough for the FP32 phase to converge to ~2e-3 at 512 bands; 64 chunks of 27,000 terms
// were 9% faster but stalled it near 7e-2.  FP64: 64 chunks (the fastest measured).  want > 0
// (the solver's gram_chunks option) chooses the count.
template <class T>
static int64_t gram_chunks(int64_t K, int want = 0) {
    if (want > 0) return std::max<int64_t>(1, std::min<int64_t>(want, K / 64));
    return std::max<int64_t>(1, std::min<int64_t>(sizeof(T) == 4 ? 256 : 64, K / 2048));
}
template <class T>
static size_t gram_workspace(int64_t K, int M, int N, int want = 0) {
    return (size_t)(gram_chunks<T>(K, want) + 1) * M * N * sizeof(T);
}

// C (M x N, FP64, column stride ldc) = A^T B over K rows; A: M columns of length K (column m
// at A + m lda), B: N columns (ld ldb).  The Gram blocks of the Rayleigh-Ritz matrices
// (MakeMatrix_d_blas, lobpcg_d_multi_mpi.h:192-250): split-K over the rows as one
// strided-batched library GEMM in the block precision (enough parallel work for small M, N),
// then the chunks summed in FP64.
template <class T>
static void gram_splitk(rocblas_handle hb, hipStream_t s, const T* A, int64_t lda, const T* B, int64_t ldb, int64_t K,
                        int M, int N, double* C, int ldc, T* ws, size_t ws_bytes, int want_chunks = 0) {
    const int64_t nch = gram_chunks<T>(K, want_chunks);
    const int64_t kc = K / nch;
    const int64_t rem = K - nch * kc;
    const int64_t nn = (int64_t)M * N;
    QG_REQUIRE(ws_bytes >= (size_t)(nch + 1) * nn * sizeof(T), "split-K workspace too small");
    if constexpr (sizeof(T) == 8) {
        const double one = 1.0, zero = 0.0;
        QG_BLAS(rocblas_dgemm_strided_batched(hb, rocblas_operation_transpose, rocblas_operation_none, M, N, (int)kc,
                                              &one, A, (int)lda, kc, B, (int)ldb, kc, &zero, ws, M, nn, (int)nch));
        if (rem)
            QG_BLAS(rocblas_dgemm(hb, rocblas_operation_transpose, rocblas_operation_none, M, N, (int)rem, &one,
                                  A + nch * kc, (int)lda, B + nch * kc, (int)ldb, &zero, ws + nch * nn, M));
    } else {
        const float one = 1.f, zero = 0.f;
        QG_BLAS(rocblas_sgemm_strided_batched(hb, rocblas_operation_transpose, rocblas_operation_none, M, N, (int)kc,
                                              &one, A, (int)lda, kc, B, (int)ldb, kc, &zero, ws, M, nn, (int)nch));
        if (rem)
            QG_BLAS(rocblas_sgemm(hb, rocblas_operation_transpose, rocblas_operation_none, M, N, (int)rem, &one,
                                  A + nch * kc, (int)lda, B + nch * kc, (int)ldb, &zero, ws + nch * nn, M));
    }
    hipLaunchKernelGGL((split_reduce_kernel<T>), dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, ws,
                       (int)(nch + (rem ? 1 : 0)), M, N, C, ldc);
}

// ------------------------------------------------------------------ solver state ----
struct LobpcgState {
    int N = 0;
    int64_t ng = 0;
    int prec = -1;
    rocblas_handle blas = nullptr;
    hipStream_t stream = nullptr;
    // vector blocks in the working precision (bytes), each N x ng
    DevBuf<unsigned char> Xw, Pw, R, AX, AR, AP, T;
    DevBuf<double> Sa, Sb;        // 3N x 3N + 4 N x N scratch each (reference layout)
    DevBuf<double> A3, B3, D, E;  // nb x nb eigenproblem
    DevBuf<int> info;
    DevBuf<double> eig, norm_a, norm_b, res2;
    DevBuf<float> C32, G32, gram_ws;
    DevBuf<unsigned char> split_ws;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~LobpcgState() {
        if (blas) rocblas_destroy_handle(blas);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

void lobpcg_state_free(LobpcgState* s) { delete s; }

// shift of the kinetic preconditioner: the run's own, or the fallback of a learned run
static double kinetic_shift(const qg_lobpcg_opts& o) {
    return (o.precond == QG_PRECOND_LEARNED && o.fallback_shift > 0.0) ? o.fallback_shift : o.precond_shift;
}

template <class T>
struct Ops {
    LobpcgState& st;
    hipStream_t s;
    int N;
    int64_t K;
    int prec;
    int nch = 0;    // split-K Gram chunks (0: the block precision's default)
    T* blk(DevBuf<unsigned char>& b) { return reinterpret_cast<T*>(b.p); }

    // C (ld ldc, FP64) = A^T B over K rows, A and B are N columns of length K (ld K).
    // Split-K: the K range is cut into chunks computed as one strided-batched library GEMM
    // (enough parallel work even for small N), then summed in FP64.
    // In QG_PREC_MIXED the H blocks (s_matrix = false) use split-BF16 MFMA; the overlap blocks
    // stay FP32, since S must remain positive definite when R and P become nearly collinear.
    void gram(const T* A, const T* B, double* C, int ldc, bool s_matrix = false) { gram_wide(A, B, N, C, ldc, s_matrix); }
    // C (N x nb) = A^T B for nb consecutive columns of B (ld K)
    void gram_wide(const T* A, const T* B, int nb, double* C, int ldc, bool s_matrix = false) {
        if constexpr (sizeof(T) == 4) {
            if (prec == QG_PREC_MIXED && !s_matrix) {
                gram_bf16(A, B, K, N, nb, K, K, C, ldc, st.gram_ws.p, st.gram_ws.n, s, true);
                return;
            }
        }
        gram_splitk<T>(st.blas, s, A, K, B, K, K, N, nb, C, ldc, reinterpret_cast<T*>(st.split_ws.p), st.split_ws.n,
                       nch);
    }
    static size_t split_workspace(int64_t K, int N, int nb_max = 0, int want = 0) {
        return gram_workspace<T>(K, N, nb_max > 0 ? nb_max : N, want);
    }
    // Out = A * Cm + beta * Out, A: K x N (ld K), Cm: N x N FP64 (ld ldc) / FP32 copy for T = float
    void update(const T* A, const double* Cm, const float* Cm32, int ldc, double beta, T* Out) {
        if constexpr (sizeof(T) == 8) {
            const double one = 1.0;
            QG_BLAS(rocblas_dgemm(st.blas, rocblas_operation_none, rocblas_operation_none, (int)K, N, N, &one, A,
                                  (int)K, Cm, ldc, &beta, Out, (int)K));
        } else {
            const float one = 1.f, b = (float)beta;
            QG_BLAS(rocblas_sgemm(st.blas, rocblas_operation_none, rocblas_operation_none, (int)K, N, N, &one, A,
                                  (int)K, Cm32, ldc, &b, Out, (int)K));
        }
    }
};

// small FP64 N x N x N products on the device: C = op(A) op(B) + beta C
static void mm(rocblas_handle h, bool ta, bool tb, int N, const double* A, int lda, const double* B, int ldb,
               double beta, double* C, int ldc) {
    const double one = 1.0;
    QG_BLAS(rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                          tb ? rocblas_operation_transpose : rocblas_operation_none, N, N, N, &one, A, lda, B, ldb,
                          &beta, C, ldc));
}

static void copy_sub(int ncol, int nrow, double* d, int ldd, const double* src, int lds, hipStream_t s) {
    QG_HIP(hipMemcpy2DAsync(d, (size_t)ldd * 8, src, (size_t)lds * 8, (size_t)nrow * 8, ncol,
                            hipMemcpyDeviceToDevice, s));
}

// NextMatrix_d (lobpcg_d_multi_mpi.h:265-460) for one of Sa / Sb: the P'P', X'P', X'X'
// blocks from the Ritz coefficients.  tmp region = 4 N x N after the 3N x 3N matrix.
static void next_matrix(rocblas_handle h, hipStream_t s, int N, double* S, const double* Cx, const double* Cr,
                        const double* Cp, int ldc, bool use_p) {
    const int n3 = 3 * N;
    const int64_t nn = (int64_t)N * N, nn9 = (int64_t)n3 * n3;
    double* XX = S;
    double* XR = S + (int64_t)N * n3;
    double* XP = S + (int64_t)N * 2 * n3;
    double* RR = S + (int64_t)N * (n3 + 1);
    double* RP = S + (int64_t)N * (2 * n3 + 1);
    double* PP = S + (int64_t)N * (2 * n3 + 2);
    double* t1 = S + nn9;
    double* tPP = t1 + nn;
    double* tXP = t1 + 2 * nn;
    double* tXX = t1 + 3 * nn;
    mm(h, false, false, N, RR, n3, Cr, ldc, 0.0, t1, N);
    if (use_p) mm(h, false, false, N, RP, n3, Cp, ldc, 1.0, t1, N);
    mm(h, true, false, N, Cr, ldc, t1, N, 0.0, tPP, N);
    if (use_p) {
        mm(h, true, false, N, RP, n3, Cr, ldc, 0.0, t1, N);
        mm(h, false, false, N, PP, n3, Cp, ldc, 1.0, t1, N);
        mm(h, true, false, N, Cp, ldc, t1, N, 1.0, tPP, N);
    }
    mm(h, false, false, N, XR, n3, Cr, ldc, 0.0, t1, N);
    if (use_p) mm(h, false, false, N, XP, n3, Cp, ldc, 1.0, t1, N);
    copy_sub(N, N, tXP, N, tPP, N, s);
    mm(h, true, false, N, Cx, ldc, t1, N, 1.0, tXP, N);
    mm(h, false, false, N, XX, n3, Cx, ldc, 1.0, t1, N);
    copy_sub(N, N, tXX, N, tXP, N, s);
    mm(h, true, false, N, t1, N, Cx, ldc, 1.0, tXX, N);
    copy_sub(N, N, PP, n3, tPP, N, s);
    copy_sub(N, N, XP, n3, tXP, N, s);
    copy_sub(N, N, XX, n3, tXX, N, s);
}

template <class T>
static void run_lobpcg(qg_ham* h, const qg_lobpcg_opts& o, int N, double* eig_host, double* X64, double* P64,
                       int scf_step, qg_lobpcg_stats& stats, int iter_offset = 0) {
    const int nconv = (o.nconv > 0 && o.nconv < N) ? o.nconv : N;   // bands the residual tests look at
    const qg_grid& g = h->grid;
    const int64_t K = g.ng();
    const double dvol = g.dvol();
    const int dtype = sizeof(T) == 8 ? QG_F64 : QG_F32;
    if (!h->lobpcg) h->lobpcg = new LobpcgState;
    LobpcgState& st = *h->lobpcg;
    if (!st.blas) {
        QG_BLAS(rocblas_create_handle(&st.blas));
        QG_HIP(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
        QG_BLAS(rocblas_set_stream(st.blas, st.stream));
        for (auto& e : st.ev) QG_HIP(hipEventCreate(&e));
    }
    hipStream_t s = st.stream;
    const size_t blk = (size_t)K * N * sizeof(T);
    st.R.alloc(blk); st.AX.alloc(blk); st.AR.alloc(blk); st.AP.alloc(blk); st.T.alloc(blk);
    if (sizeof(T) != 8) { st.Xw.alloc(blk); st.Pw.alloc(blk); }
    const int n3 = 3 * N;
    const int64_t nn9 = (int64_t)n3 * n3;
    st.Sa.alloc(nn9 + 4 * (int64_t)N * N);
    st.Sb.alloc(nn9 + 4 * (int64_t)N * N);
    st.A3.alloc(nn9); st.B3.alloc(nn9); st.D.alloc(n3); st.E.alloc(n3); st.info.alloc(1);
    st.eig.alloc(N); st.norm_a.alloc(N); st.norm_b.alloc(N); st.res2.alloc(N);
    if (sizeof(T) != 8) { st.C32.alloc(nn9); st.G32.alloc((size_t)N * N); }
    if (o.precision == QG_PREC_MIXED) st.gram_ws.alloc(gram_bf16_workspace(K, N, N));
    st.split_ws.alloc(Ops<T>::split_workspace(K, N, 0, o.gram_chunks));
    QG_HIP(hipMemsetAsync(st.Sa.p, 0, (nn9 + 4 * (int64_t)N * N) * 8, s));
    QG_HIP(hipMemsetAsync(st.Sb.p, 0, (nn9 + 4 * (int64_t)N * N) * 8, s));

    Ops<T> op{st, s, N, K, o.precision, o.gram_chunks};
    T* X = (sizeof(T) == 8) ? reinterpret_cast<T*>(X64) : op.blk(st.Xw);
    T* P = (sizeof(T) == 8) ? reinterpret_cast<T*>(P64) : op.blk(st.Pw);
    T* R = op.blk(st.R);
    T* AX = op.blk(st.AX);
    T* AR = op.blk(st.AR);
    T* AP = op.blk(st.AP);
    T* Tm = op.blk(st.T);
    const int64_t KN = K * N;
    const int gsz = grid_for(KN);
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
    // sector clock: "others" up to the start of a sector, then the sector itself
    // (the reference's WATCH::Record ids)
    auto sector_begin = [&] {
        if (o.watch) o.watch(o.watch_user, 10, 1);
        return now();
    };
    auto sector_end = [&](std::chrono::steady_clock::time_point t0, int idx, int64_t count) {
        const double d = secs(t0, now());
        if (o.watch) o.watch(o.watch_user, idx, count);
        return d;
    };
    auto sync = [&] { QG_HIP(hipStreamSynchronize(s)); };
    auto hpsi = [&](const T* in, T* outp) {
        auto t0 = sector_begin();
        if (o.op)
            o.op(o.op_user, dtype, N, in, outp, (void*)s);
        else if (h->batch)
            batch_service_apply(h->batch, h, dtype, N, in, outp, s);
        else
            hamiltonian_apply(h, dtype, N, in, K, outp, K, s);
        sync();
        stats.t_hpsi += sector_end(t0, 11, N);
        stats.hpsi_calls += N;
    };

    // entry: X scaled by sqrt(dVol) (:954-956); working copies for FP32 modes
    const double sq = std::sqrt(dvol);
    if (sizeof(T) == 8) {
        hipLaunchKernelGGL((scale_kernel<double>), dim3(gsz), dim3(256), 0, s, X64, KN, sq);
    } else {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(gsz), dim3(256), 0, s, X64, (float*)X, KN, sq);
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(gsz), dim3(256), 0, s, P64, (float*)P, KN, 1.0);
    }
    const bool first_time = (scf_step == 0);
    if (first_time) QG_HIP(hipMemsetAsync(P, 0, blk, s));
    hpsi(X, AX);
    if (first_time) {
        QG_HIP(hipMemsetAsync(AP, 0, blk, s));
    } else {
        hpsi(P, AP);
    }
    hipLaunchKernelGGL((coldot_kernel<T>), dim3(N), dim3(256), 0, s, X, AX, K, K, st.eig.p);
    hipLaunchKernelGGL((residual_kernel<T>), dim3(N), dim3(256), 0, s, AX, X, st.eig.p, R, K, K, st.res2.p);
    std::vector<double> ev(n3), nx(N), np(N), r2(N);
    QG_HIP(hipMemcpyAsync(eig_host, st.eig.p, N * 8, hipMemcpyDeviceToHost, s));
    QG_HIP(hipMemcpyAsync(r2.data(), st.res2.p, N * 8, hipMemcpyDeviceToHost, s));
    sync();
    double maxres = 0.0;
    for (int k = 0; k < nconv; ++k) maxres = std::max(maxres, std::sqrt(r2[k]));
    stats.max_residual = maxres;

    const bool precond = o.precond != QG_PRECOND_NONE;
    // FP32 working precision: columns of R, P (and AP) are normalised before every Gram
    // product and X is renormalised after every update, all blocks are recomputed each
    // iteration, and a norm drift is absorbed instead of rolling back (the reference's
    // recurrences and rollback assume FP64 arithmetic).
    const bool robust = sizeof(T) != 8;
    // With a preconditioner (an extension of the reference) a norm drift or a failed
    // generalized eigensolve restarts the iteration in place instead of leaving it: the step
    // is discarded, P is dropped (the next basis is [X R]) and the residual is recomputed.
    // The reference's exit with P = R would put K R twice into the basis of a following call
    // on the same operator.  The unpreconditioned run keeps the reference's rollback exactly.
    const bool soft_restart = precond || robust;
    bool p_valid = !first_time;
    const bool direct_xr = !precond && !robust;   // DIRECT_S_MATRIX_XR_RP needs R = AX - XE
    auto blkS = [&](DevBuf<double>& S, int bi, int bj) { return S.p + (int64_t)N * bi + (int64_t)N * n3 * bj; };
    bool large_diff = false;
    int iter = 0;
    bool finished = false;
    // learned preconditioner: replaced by the kinetic one when the block's residual norm
    // (sqrt of the sum over bands; the largest single residual is set by the slowest band at
    // the block's edge and says little about the preconditioner) has not reached a new minimum
    // for stall_window iterations
    bool learned = o.precond == QG_PRECOND_LEARNED && h->cnn && stats.fallback_iter < 0;
    double best_res = 1.0e300;
    int best_iter = 0;
    for (iter = 0; iter < o.max_iter && !finished; ++iter) {
        auto t0 = sector_begin();
        if (precond) {
            if (learned)
                cnn_precond_apply(h->cnn, dtype, N, R, K, s);
            else
                precond_apply(h, QG_PRECOND_KINETIC, kinetic_shift(o), dtype, N, R, K, s);
            sync();
            stats.t_precond += sector_end(t0, 10, 1);
        }
        const bool use_p = p_valid;
        if (robust) {
            hipLaunchKernelGGL((coldot_kernel<T>), dim3(N), dim3(256), 0, s, R, R, K, K, st.norm_a.p);
            hipLaunchKernelGGL((colscale_kernel<T>), dim3(grid_for(K) / 4 + 1, N), dim3(256), 0, s, R, (T*)nullptr,
                               st.norm_a.p, K, K);
            if (use_p) {
                hipLaunchKernelGGL((coldot_kernel<T>), dim3(N), dim3(256), 0, s, P, P, K, K, st.norm_b.p);
                hipLaunchKernelGGL((colscale_kernel<T>), dim3(grid_for(K) / 4 + 1, N), dim3(256), 0, s, P, AP,
                                   st.norm_b.p, K, K);
            }
        }
        hpsi(R, AR);
        // MakeMatrix_d_blas (:192-250)
        t0 = sector_begin();
        const bool full = robust || (iter == 0) || (o.refresh_period > 0 && iter % o.refresh_period == 0) || large_diff;
        if (full) {
            op.gram(X, AX, blkS(st.Sa, 0, 0), n3);
            op.gram(X, AP, blkS(st.Sa, 0, 2), n3);
            op.gram(P, AP, blkS(st.Sa, 2, 2), n3);
            op.gram(X, X, blkS(st.Sb, 0, 0), n3, true);
            op.gram(X, P, blkS(st.Sb, 0, 2), n3, true);
            op.gram(P, P, blkS(st.Sb, 2, 2), n3, true);
            large_diff = false;
        }
        if (full || !direct_xr) {
            op.gram(X, R, blkS(st.Sb, 0, 1), n3, true);
            op.gram(R, P, blkS(st.Sb, 1, 2), n3, true);
        }
        op.gram(X, AR, blkS(st.Sa, 0, 1), n3);
        op.gram(R, AR, blkS(st.Sa, 1, 1), n3);
        op.gram(R, AP, blkS(st.Sa, 1, 2), n3);
        op.gram(R, R, blkS(st.Sb, 1, 1), n3, true);
        sync();
        stats.t_gram += sector_end(t0, 17, 1);

        // Rayleigh-Ritz: dsygvd on the upper nb x nb problem (lapack_DSYGVX, w_dsygvx.h:25-103)
        t0 = sector_begin();
        const int nb = use_p ? n3 : 2 * N;
        copy_sub(nb, nb, st.A3.p, nb, st.Sa.p, n3, s);
        copy_sub(nb, nb, st.B3.p, nb, st.Sb.p, n3, s);
        QG_BLAS(rocsolver_dsygvd(st.blas, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, nb, st.A3.p,
                                 nb, st.B3.p, nb, st.D.p, st.E.p, st.info.p));
        int info = 0;
        QG_HIP(hipMemcpyAsync(&info, st.info.p, sizeof(int), hipMemcpyDeviceToHost, s));
        QG_HIP(hipMemcpyAsync(ev.data(), st.D.p, nb * 8, hipMemcpyDeviceToHost, s));
        sync();
        stats.t_eig += sector_end(t0, 15, 1);
        if (info != 0 && o.verbose)
            std::fprintf(stderr, "lobpcg: dsygvd info = %d at iteration %d%s\n", info, iter,
                         soft_restart ? " (restart without P)" : " (ignored, as the reference)");
        auto restart = [&] {
            // discard this step: X and AX stay, P and AP are dropped, R = AX - X E again
            QG_HIP(hipMemsetAsync(P, 0, blk, s));
            QG_HIP(hipMemsetAsync(AP, 0, blk, s));
            hipLaunchKernelGGL((residual_kernel<T>), dim3(N), dim3(256), 0, s, AX, X, st.eig.p, R, K, K, st.res2.p);
            p_valid = false;
            large_diff = true;   // full Gram on the next iteration
            ++stats.rollbacks;
            sync();
        };
        if (info != 0 && soft_restart) {
            restart();
            continue;
        }

        const double* Cx = st.A3.p;
        const double* Cr = st.A3.p + N;
        const double* Cp = st.A3.p + 2 * N;
        const float* Cx32 = nullptr;
        const float* Cr32 = nullptr;
        const float* Cp32 = nullptr;
        if (sizeof(T) != 8) {
            hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for((int64_t)nb * nb)), dim3(256), 0, s,
                               st.A3.p, st.C32.p, (int64_t)nb * nb, 1.0);
            Cx32 = st.C32.p; Cr32 = st.C32.p + N; Cp32 = st.C32.p + 2 * N;
        }
        const bool final_step = (iter + 1 == o.max_iter);
        // NextVectors2(x, r, p, tmp): T = R Cr (+ P Cp); P = T; T += X Cx
        t0 = sector_begin();
        op.update(R, Cr, Cr32, nb, 0.0, Tm);
        if (use_p) op.update(P, Cp, Cp32, nb, 1.0, Tm);
        QG_HIP(hipMemcpyAsync(P, Tm, blk, hipMemcpyDeviceToDevice, s));
        op.update(X, Cx, Cx32, nb, 1.0, Tm);
        hipLaunchKernelGGL((coldot_kernel<T>), dim3(N), dim3(256), 0, s, Tm, Tm, K, K, st.norm_a.p);
        QG_HIP(hipMemcpyAsync(nx.data(), st.norm_a.p, N * 8, hipMemcpyDeviceToHost, s));
        sync();
        stats.t_update += sector_end(t0, 16, 1);
        double maxd = 0.0;
        for (int k = 0; k < N; ++k) maxd = std::max(maxd, std::fabs(1.0 - nx[k]));
        large_diff = !robust && maxd > o.rollback_limit;
        if (large_diff && soft_restart) {
            if (o.verbose) std::fprintf(stderr, "lobpcg: |1-|x|^2| = %g at iteration %d, restart without P\n", maxd, iter);
            restart();
            continue;
        }
        if (large_diff) {
            // rollback (:1322-1346): P = R, X keeps the previous vectors, leave the loop
            QG_HIP(hipMemcpyAsync(P, R, blk, hipMemcpyDeviceToDevice, s));
            ++stats.rollbacks;
            if (o.verbose) std::fprintf(stderr, "lobpcg: rollback |1-|x|^2| = %g at iteration %d\n", maxd, iter);
            finished = true;
            // X back to the grid normalisation
            if (sizeof(T) == 8)
                hipLaunchKernelGGL((scale_kernel<double>), dim3(gsz), dim3(256), 0, s, X64, KN, 1.0 / sq);
            else
                hipLaunchKernelGGL((convert_kernel<float, double>), dim3(gsz), dim3(256), 0, s, (const float*)X, X64,
                                   KN, 1.0 / sq);
            break;
        }
        QG_HIP(hipMemcpyAsync(X, Tm, blk, hipMemcpyDeviceToDevice, s));
        p_valid = true;
        for (int i = 0; i < N; ++i) eig_host[i] = ev[i];
        QG_HIP(hipMemcpyAsync(st.eig.p, st.D.p, N * 8, hipMemcpyDeviceToDevice, s));
        bool stop = final_step;
        if (!final_step) {
            t0 = sector_begin();
            // NextVectors2(Ax, Ar, Ap, tmp); AX = T
            op.update(AR, Cr, Cr32, nb, 0.0, Tm);
            if (use_p) op.update(AP, Cp, Cp32, nb, 1.0, Tm);
            QG_HIP(hipMemcpyAsync(AP, Tm, blk, hipMemcpyDeviceToDevice, s));
            op.update(AX, Cx, Cx32, nb, 1.0, Tm);
            QG_HIP(hipMemcpyAsync(AX, Tm, blk, hipMemcpyDeviceToDevice, s));
            if (robust)   // renormalise X (and AX with it); norm_a holds ||x'||^2
                hipLaunchKernelGGL((colscale_kernel<T>), dim3(grid_for(K) / 4 + 1, N), dim3(256), 0, s, X, AX,
                                   st.norm_a.p, K, K);
            hipLaunchKernelGGL((residual_kernel<T>), dim3(N), dim3(256), 0, s, AX, X, st.eig.p, R, K, K, st.res2.p);
            sync();
            stats.t_update += sector_end(t0, 16, 1);
            t0 = sector_begin();
            // NextMatrix_d: recurrences for the S and H blocks, then XtR / RtP directly
            next_matrix(st.blas, s, N, st.Sb.p, Cx, Cr, Cp, nb, use_p);
            next_matrix(st.blas, s, N, st.Sa.p, Cx, Cr, Cp, nb, use_p);
            if (direct_xr) {
                const int64_t nn = (int64_t)N * N;
                hipLaunchKernelGGL(direct_xr_rp_kernel, dim3((N + 255) / 256, N), dim3(256), 0, s, N,
                                   blkS(st.Sb, 0, 1), blkS(st.Sb, 1, 2), n3, st.Sa.p + nn9 + 3 * nn,
                                   st.Sb.p + nn9 + 3 * nn, st.Sa.p + nn9 + 2 * nn, st.Sb.p + nn9 + 2 * nn,
                                   st.eig.p);
            }
            QG_HIP(hipMemcpyAsync(r2.data(), st.res2.p, N * 8, hipMemcpyDeviceToHost, s));
            sync();
            stats.t_other += sector_end(t0, 10, 1);
            maxres = 0.0;
            double totres = 0.0;
            for (int k = 0; k < nconv; ++k) {
                maxres = std::max(maxres, std::sqrt(r2[k]));
                totres += r2[k];
            }
            totres = std::sqrt(totres);
            stats.max_residual = maxres;
            if (o.verbose >= 2)
                std::fprintf(stderr, "lobpcg it %d: e0 %.10f e%d %.10f max|r| %.3e\n", iter, ev[0], N - 1, ev[N - 1],
                             maxres);
            if (o.tol > 0.0 && maxres <= o.tol) stop = true;
            if (totres < best_res) {
                best_res = totres;
                best_iter = iter;
            } else if (learned && o.stall_window > 0 && iter - best_iter >= o.stall_window) {
                learned = false;
                stats.fallback_iter = iter_offset + iter + 1;
                if (o.verbose) std::fprintf(stderr, "lobpcg: learned preconditioner stalled, kinetic from iteration %d\n",
                                            stats.fallback_iter);
            }
        }
        if (stop) {
            // final normalisation x_j /= sqrt(||x_j||^2 dVol) (:1391-1394)
            std::vector<double> c(N);
            for (int j = 0; j < N; ++j) c[j] = 1.0 / std::sqrt((robust && !final_step ? 1.0 : nx[j]) * dvol);
            DevBuf<double> dc;
            dc.upload(c.data(), N, s);
            for (int j = 0; j < N; ++j) {
                if (sizeof(T) == 8)
                    hipLaunchKernelGGL((scale_kernel<double>), dim3(grid_for(K)), dim3(256), 0, s, X64 + (int64_t)j * K,
                                       K, c[j]);
                else
                    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(K)), dim3(256), 0, s,
                                       (const float*)X + (int64_t)j * K, X64 + (int64_t)j * K, K, c[j]);
            }
            finished = true;
            ++iter;
            break;
        }
    }
    if (!finished) {
        // max_iter == 0: undo the entry scaling
        if (sizeof(T) == 8)
            hipLaunchKernelGGL((scale_kernel<double>), dim3(gsz), dim3(256), 0, s, X64, KN, 1.0 / sq);
        else
            hipLaunchKernelGGL((convert_kernel<float, double>), dim3(gsz), dim3(256), 0, s, (const float*)X, X64, KN,
                               1.0 / sq);
    }
    if (sizeof(T) != 8)
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(gsz), dim3(256), 0, s, (const float*)P, P64, KN, 1.0);
    sync();
    stats.iterations = iter;
}


// ------------------------------------------------------- orthonormal-basis LOBPCG ----
// The iteration used with a preconditioner or in FP32: the Rayleigh-Ritz basis [X W P] is kept
// orthonormal (Hetmaniuk and Lehoucq's basis selection), so the small problem is a standard
// symmetric eigenproblem and near-dependence of W or P with X cannot make it indefinite:
//   W = K R;  W -= X (X^T W);  W = SVQB(W)
//   P -= X (X^T P) + W (W^T P);  P = SVQB(P)          (AP follows with the same coefficients)
//   [X W P]^T H [X W P] c = theta c  (lowest N)  ->  X', P' = W Cw + P Cp, AX', AP'
// SVQB(V): G = V^T V, D = diag(G)^-1/2, D G D = U diag(theta) U^T, V <- V D U theta^-1/2, with
// the directions of theta <= tau theta_max dropped (zero columns, excluded from the small
// problem by a large diagonal).  The Ritz values and the residual definition are the
// reference's (lobpcg_d_multi_mpi.h:1280-1320).

// d_i = G_ii^-1/2 (0 for an empty column)
__global__ void svqb_diag_kernel(const double* G, int N, int ld, double* d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double gi = G[i + (int64_t)ld * i];
    d[i] = gi > 0.0 ? 1.0 / sqrt(gi) : 0.0;
}

// G <- D G D (d computed beforehand: the diagonal is rewritten here)
__global__ void svqb_scale_kernel(double* G, int N, int ld, const double* d) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    G[i + (int64_t)ld * j] *= d[i] * d[j];
}

// T_ij = d_i U_ij theta_j^-1/2 for kept directions, 0 otherwise; keep_j = 1 / 0
__global__ void svqb_transform_kernel(const double* U, int N, int ld, const double* theta, const double* d,
                                      double tau, double* Tm, int ldt, int* keep) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double tmax = theta[N - 1];
    const bool k = theta[j] > tau * tmax && theta[j] > 0.0;
    Tm[i + (int64_t)ldt * j] = k ? d[i] * U[i + (int64_t)ld * j] / sqrt(theta[j]) : 0.0;
    if (i == 0) keep[j] = k ? 1 : 0;
}

// Cholesky-QR transform T = D R^-1: T_ij = d_i (R^-1)_ij for i <= j and 0 below the diagonal
// (R^-1 is in the upper triangle of Rinv, whose strict lower part holds stale values); keep_j = 1
__global__ void chol_transform_kernel(const double* Rinv, int N, int ld, const double* d, double* Tm, int ldt,
                                      int* keep) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    Tm[i + (int64_t)ldt * j] = (i <= j) ? d[i] * Rinv[i + (int64_t)ld * j] : 0.0;
    if (i == 0) keep[j] = 1;
}

// pivots of a Cholesky factor: p_j = R_jj^2
__global__ void chol_pivots_kernel(const double* R, int N, int ld, double* p) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const double r = R[j + (int64_t)ld * j];
    p[j] = r * r;
}

// dropped basis directions: large diagonal in the Rayleigh-Ritz matrix, so that none of the
// lowest N eigenpairs uses them
__global__ void mask_diag_kernel(double* A, int ld, const int* keep, int off, int N, double big) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N || keep[j]) return;
    A[(off + j) + (int64_t)ld * (off + j)] = big;
}

// copy an m x n FP64 matrix (ld) into a contiguous FP32 one with an optional sign
__global__ void narrow_kernel(const double* A, int m, int n, int ld, float* B, double sgn) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)m * n) return;
    B[i] = (float)(sgn * A[(i % m) + (int64_t)ld * (i / m)]);
}

__global__ void diag_copy_kernel(const double* D, double* e, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) e[i] = D[i];
}

template <class T>
static void run_lobpcg_ortho(qg_ham* h, const qg_lobpcg_opts& o, int N, double* eig_host, double* X64, double* P64,
                             int scf_step, qg_lobpcg_stats& stats, int iter_offset = 0, int switch_stall = 0) {
    const int nconv = (o.nconv > 0 && o.nconv < N) ? o.nconv : N;   // bands the residual tests look at
    const qg_grid& g = h->grid;
    const int64_t K = g.ng();
    const double dvol = g.dvol();
    const int dtype = sizeof(T) == 8 ? QG_F64 : QG_F32;
    if (!h->lobpcg) h->lobpcg = new LobpcgState;
    LobpcgState& st = *h->lobpcg;
    if (!st.blas) {
        QG_BLAS(rocblas_create_handle(&st.blas));
        QG_HIP(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
        QG_BLAS(rocblas_set_stream(st.blas, st.stream));
        for (auto& e : st.ev) QG_HIP(hipEventCreate(&e));
    }
    hipStream_t s = st.stream;
    const size_t blk = (size_t)K * N * sizeof(T);
    // AX, AW, AP are consecutive blocks of one allocation: the Rayleigh-Ritz Grams are three
    // wide products X^T [AX AW AP], W^T [AW AP], P^T AP
    st.R.alloc(blk); st.AX.alloc(3 * blk); st.T.alloc(blk);
    if (sizeof(T) != 8) { st.Xw.alloc(blk); st.Pw.alloc(blk); }
    const int n3 = 3 * N;
    const int64_t nn = (int64_t)N * N, nn9 = (int64_t)n3 * n3;
    st.A3.alloc(nn9); st.D.alloc(n3); st.E.alloc(n3); st.info.alloc(1);
    st.eig.alloc(N); st.norm_a.alloc(N); st.norm_b.alloc(N); st.res2.alloc(N);
    st.C32.alloc(nn9);
    if (o.precision == QG_PREC_MIXED) st.gram_ws.alloc(gram_bf16_workspace(K, N, n3));
    st.split_ws.alloc(Ops<T>::split_workspace(K, N, n3, o.gram_chunks));
    DevBuf<double> Gm, Tr, dsc, Rc;
    Gm.alloc(nn); Tr.alloc(nn); dsc.alloc(N); Rc.alloc(nn);
    DevBuf<int> keep;
    keep.alloc(n3);

    Ops<T> op{st, s, N, K, o.precision, o.gram_chunks};
    // X and P live in the caller's FP64 arrays for T = double; blocks are swapped by pointer
    T* X = (sizeof(T) == 8) ? reinterpret_cast<T*>(X64) : op.blk(st.Xw);
    T* P = (sizeof(T) == 8) ? reinterpret_cast<T*>(P64) : op.blk(st.Pw);
    T* W = op.blk(st.R);
    T* const AX = op.blk(st.AX);
    T* const AW = AX + (int64_t)K * N;
    T* const AP = AX + 2 * (int64_t)K * N;
    T* Tm = op.blk(st.T);
    const int64_t KN = K * N;
    const int gsz = grid_for(KN);
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
    // sector clock: "others" up to the start of a sector, then the sector itself
    // (the reference's WATCH::Record ids)
    auto sector_begin = [&] {
        if (o.watch) o.watch(o.watch_user, 10, 1);
        return now();
    };
    auto sector_end = [&](std::chrono::steady_clock::time_point t0, int idx, int64_t count) {
        const double d = secs(t0, now());
        if (o.watch) o.watch(o.watch_user, idx, count);
        return d;
    };
    auto sync = [&] { QG_HIP(hipStreamSynchronize(s)); };
    auto hpsi = [&](const T* in, T* outp) {
        auto t0 = sector_begin();
        if (o.op)
            o.op(o.op_user, dtype, N, in, outp, (void*)s);
        else if (h->batch)
            batch_service_apply(h->batch, h, dtype, N, in, outp, s);
        else
            hamiltonian_apply(h, dtype, N, in, K, outp, K, s);
        sync();
        stats.t_hpsi += sector_end(t0, 11, N);
        stats.hpsi_calls += N;
    };
    // Out = alpha A C + beta Out, C: N x N FP64 (ld ldc), A: K x N
    auto upd = [&](const T* A, const double* C, int ldc, double alpha, double beta, T* Out) {
        if constexpr (sizeof(T) == 8) {
            QG_BLAS(rocblas_dgemm(st.blas, rocblas_operation_none, rocblas_operation_none, (int)K, N, N, &alpha, A,
                                  (int)K, C, ldc, &beta, Out, (int)K));
        } else {
            hipLaunchKernelGGL(narrow_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, C, N, N, ldc,
                               st.C32.p, 1.0);
            const float a = (float)alpha, b = (float)beta;
            QG_BLAS(rocblas_sgemm(st.blas, rocblas_operation_none, rocblas_operation_none, (int)K, N, N, &a, A,
                                  (int)K, st.C32.p, N, &b, Out, (int)K));
        }
    };
    auto eigh = [&](double* A, int n, int lda, double* w) -> int {
        auto t0 = sector_begin();
        QG_BLAS(rocsolver_dsyevd(st.blas, rocblas_evect_original, rocblas_fill_upper, n, A, lda, w, st.E.p, st.info.p));
        int info = 0;
        QG_HIP(hipMemcpyAsync(&info, st.info.p, sizeof(int), hipMemcpyDeviceToHost, s));
        sync();
        stats.t_eig += sector_end(t0, 15, 1);
        return info;
    };
    // drop threshold of SVQB: well above the relative rounding of the Gram matrix (FP32 Grams
    // carry ~1e-6), so the kept directions are amplified at most ~1/sqrt(tau)
    const double tau = sizeof(T) == 8 ? 1e-12 : 1e-4;
    // V <- SVQB(V); keep flags into keep_out.  Returns the smallest kept theta over the
    // largest (1: already orthonormal; small: the transform amplified rounding, so callers
    // orthogonalise once more).
    std::vector<double> theta(N);
    int info2[2] = {0, 0};
    auto svqb = [&](T*& V, int* keep_out, bool drop) -> double {
        auto t0 = sector_begin();
        op.gram(V, V, Gm.p, N, true);
        hipLaunchKernelGGL(svqb_diag_kernel, dim3((N + 255) / 256), dim3(256), 0, s, Gm.p, N, N, dsc.p);
        hipLaunchKernelGGL(svqb_scale_kernel, dim3((N + 255) / 256, N), dim3(256), 0, s, Gm.p, N, N,
                           (const double*)dsc.p);
        sync();
        stats.t_gram += sector_end(t0, 17, 1);
        if (o.orth != QG_ORTH_SVQB) {
            // Cholesky-QR: D G D = R^T R, V <- V D R^-1 (a triangular factor and its inverse
            // instead of an eigensolve).  A failed factor or a pivot ratio p_min / p_max at or
            // below the drop threshold (near-dependent columns) goes to SVQB, which drops the
            // dependent directions.  Returns the pivot ratio (1 for orthonormal columns).
            t0 = sector_begin();
            QG_HIP(hipMemcpyAsync(Rc.p, Gm.p, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
            QG_BLAS(rocsolver_dpotrf(st.blas, rocblas_fill_upper, N, Rc.p, N, st.info.p));
            hipLaunchKernelGGL(chol_pivots_kernel, dim3((N + 255) / 256), dim3(256), 0, s, Rc.p, N, N, st.D.p);
            QG_HIP(hipMemcpyAsync(&info2[0], st.info.p, sizeof(int), hipMemcpyDeviceToHost, s));
            QG_HIP(hipMemcpyAsync(theta.data(), st.D.p, N * 8, hipMemcpyDeviceToHost, s));
            sync();
            double pmax = 0.0, pmin = 0.0;
            if (info2[0] == 0) {
                pmax = *std::max_element(theta.begin(), theta.end());
                pmin = *std::min_element(theta.begin(), theta.end());
            }
            const double ratio = pmax > 0.0 ? pmin / pmax : 0.0;
            if (info2[0] == 0 && pmin > 0.0 && ratio > (drop ? tau : 0.0)) {
                QG_BLAS(rocsolver_dtrtri(st.blas, rocblas_fill_upper, rocblas_diagonal_non_unit, N, Rc.p, N,
                                         st.info.p));
                QG_HIP(hipMemcpyAsync(&info2[1], st.info.p, sizeof(int), hipMemcpyDeviceToHost, s));
                sync();
                if (info2[1] == 0) {
                    hipLaunchKernelGGL(chol_transform_kernel, dim3((N + 255) / 256, N), dim3(256), 0, s, Rc.p, N, N,
                                       (const double*)dsc.p, Tr.p, N, keep_out);
                    stats.t_eig += sector_end(t0, 15, 1);
                    t0 = sector_begin();
                    upd(V, Tr.p, N, 1.0, 0.0, Tm);
                    std::swap(V, Tm);
                    sync();
                    stats.t_update += sector_end(t0, 16, 1);
                    return ratio;
                }
            }
            stats.t_eig += sector_end(t0, 15, 1);
        }
        const int info = eigh(Gm.p, N, N, st.D.p);
        QG_REQUIRE(info == 0, "SVQB eigensolve failed");
        hipLaunchKernelGGL(svqb_transform_kernel, dim3((N + 255) / 256, N), dim3(256), 0, s, Gm.p, N, N, st.D.p,
                           dsc.p, drop ? tau : 0.0, Tr.p, N, keep_out);
        QG_HIP(hipMemcpyAsync(theta.data(), st.D.p, N * 8, hipMemcpyDeviceToHost, s));
        t0 = sector_begin();
        upd(V, Tr.p, N, 1.0, 0.0, Tm);
        std::swap(V, Tm);
        sync();
        stats.t_update += sector_end(t0, 16, 1);
        const double tmax = theta[N - 1];
        double tmin = tmax;
        for (int j = 0; j < N; ++j)
            if (theta[j] > (drop ? tau : 0.0) * tmax && theta[j] > 0.0) tmin = std::min(tmin, theta[j]);
        return tmax > 0.0 ? tmin / tmax : 1.0;
    };
    // second pass when the first one amplified rounding (ratio at or below 1e-3) in FP64, always
    // in FP32 (a ratio never exceeds 1); the reortho option overrides
    const double reortho = o.reortho > 0.0 ? o.reortho : (sizeof(T) == 8 ? 1e-3 : 2.0);
    // Rayleigh-Ritz on the (orthonormal) columns of V with AV: V, AV rotated to Ritz vectors
    auto rr_single = [&](T*& V, T* AV) {
        op.gram(V, AV, st.A3.p, N);
        const int info = eigh(st.A3.p, N, N, st.D.p);
        QG_REQUIRE(info == 0, "Rayleigh-Ritz eigensolve failed");
        upd(V, st.A3.p, N, 1.0, 0.0, Tm);
        std::swap(V, Tm);
        upd(AV, st.A3.p, N, 1.0, 0.0, Tm);
        QG_HIP(hipMemcpyAsync(AV, Tm, blk, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(diag_copy_kernel, dim3((N + 255) / 256), dim3(256), 0, s, st.D.p, st.eig.p, N);
    };
    // largest band residual, and the block's residual norm sqrt(sum over bands)
    double totres = 0.0;
    auto residual = [&](double& maxres) {
        hipLaunchKernelGGL((residual_kernel<T>), dim3(N), dim3(256), 0, s, AX, X, st.eig.p, W, K, K, st.res2.p);
        std::vector<double> r2(N);
        QG_HIP(hipMemcpyAsync(r2.data(), st.res2.p, N * 8, hipMemcpyDeviceToHost, s));
        QG_HIP(hipMemcpyAsync(eig_host, st.eig.p, N * 8, hipMemcpyDeviceToHost, s));
        sync();
        maxres = 0.0;
        totres = 0.0;
        for (int k = 0; k < nconv; ++k) {
            maxres = std::max(maxres, std::sqrt(r2[k]));
            totres += r2[k];
        }
        totres = std::sqrt(totres);
    };

    // entry: X scaled by sqrt(dVol) (lobpcg_d_multi_mpi.h:954-956), made orthonormal and
    // rotated to Ritz vectors
    const double sq = std::sqrt(dvol);
    if (sizeof(T) == 8) {
        hipLaunchKernelGGL((scale_kernel<double>), dim3(gsz), dim3(256), 0, s, X64, KN, sq);
    } else {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(gsz), dim3(256), 0, s, X64, (float*)X, KN, sq);
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(gsz), dim3(256), 0, s, P64, (float*)P, KN, sq);
    }
    svqb(X, keep.p, false);
    hpsi(X, AX);
    rr_single(X, AX);
    bool p_valid = scf_step != 0;   // the kept P block (AP is formed after its orthogonalisation)
    if (p_valid && sizeof(T) == 8)
        hipLaunchKernelGGL((scale_kernel<double>), dim3(gsz), dim3(256), 0, s, P64, KN, sq);
    double maxres = 0.0;
    residual(maxres);
    stats.max_residual = maxres;

    // learned preconditioner: replaced by the kinetic one when the block's residual norm has not
    // reached a new minimum for stall_window iterations (the largest band residual is set by the
    // slowest band at the block's edge)
    bool learned = o.precond == QG_PRECOND_LEARNED && h->cnn && stats.fallback_iter < 0;
    double best_res = totres;
    int best_iter = 0;
    // precision floor: the residual no longer falls by 30% within switch_stall iterations
    double floor_res = maxres;
    int floor_iter = 0;
    int iter = 0;
    for (iter = 0; iter < o.max_iter; ++iter) {
        // diagonal given to dropped basis directions: above every Ritz value of the kept
        // basis, but small enough that the eigensolver's absolute error (~eps |A|) stays below
        // the spectrum's resolution
        double emax = 0.0;
        for (int k = 0; k < N; ++k) emax = std::max(emax, std::fabs(eig_host[k]));
        const double big = 10.0 * (emax + 1.0e3);
        if (o.tol > 0.0 && maxres <= o.tol) break;
        // W = K R, orthogonal to X, orthonormal
        auto t0 = sector_begin();
        if (o.precond != QG_PRECOND_NONE) {
            if (learned) cnn_precond_apply(h->cnn, dtype, N, W, K, s);
            else precond_apply(h, QG_PRECOND_KINETIC, kinetic_shift(o), dtype, N, W, K, s);
            sync();
            stats.t_precond += sector_end(t0, 10, 1);
        }
        for (int pass = 0; pass < 2; ++pass) {
            t0 = sector_begin();
            op.gram(X, W, Gm.p, N, true);
            upd(X, Gm.p, N, -1.0, 1.0, W);
            sync();
            stats.t_gram += sector_end(t0, 17, 1);
            if (svqb(W, keep.p + N, true) > reortho) break;
        }
        hpsi(W, AW);
        if (p_valid) {
            // P against X and W, orthonormal; AP recomputed (a transformed AP would carry the
            // SVQB amplification of its rounding into the Rayleigh-Ritz matrix)
            for (int pass = 0; pass < 2; ++pass) {
                t0 = sector_begin();
                op.gram(X, P, Gm.p, N, true);
                upd(X, Gm.p, N, -1.0, 1.0, P);
                op.gram(W, P, Gm.p, N, true);
                upd(W, Gm.p, N, -1.0, 1.0, P);
                sync();
                stats.t_gram += sector_end(t0, 17, 1);
                if (svqb(P, keep.p + 2 * N, true) > reortho) break;
            }
            hpsi(P, AP);
        }
        // Rayleigh-Ritz: upper blocks of [X W P]^T [AX AW AP]
        t0 = sector_begin();
        const int nb = p_valid ? n3 : 2 * N;
        double* A = st.A3.p;
        auto blkA = [&](int bi, int bj) { return A + (int64_t)N * bi + (int64_t)nb * N * bj; };
        op.gram_wide(X, AX, nb, blkA(0, 0), nb);            // X^T [AX AW (AP)]
        op.gram_wide(W, AW, nb - N, blkA(1, 1), nb);        // W^T [AW (AP)]
        if (p_valid) op.gram(P, AP, blkA(2, 2), nb);        // P^T AP
        hipLaunchKernelGGL(mask_diag_kernel, dim3((2 * N + 255) / 256), dim3(256), 0, s, A, nb, keep.p + N, N, N, big);
        if (p_valid)
            hipLaunchKernelGGL(mask_diag_kernel, dim3((N + 255) / 256), dim3(256), 0, s, A, nb, keep.p + 2 * N, 2 * N,
                               N, big);
        sync();
        stats.t_gram += sector_end(t0, 17, 1);
        const int info = eigh(A, nb, nb, st.D.p);
        if (info != 0) {
            // drop P and retry the iteration from X
            if (o.verbose) std::fprintf(stderr, "lobpcg: Rayleigh-Ritz info %d at iteration %d, restart without P\n", info, iter);
            ++stats.rollbacks;
            p_valid = false;
            residual(maxres);
            continue;
        }
        // X' = X Cx + W Cw + P Cp,  P' = W Cw + P Cp; AX' = H X' (one H psi costs less than the
        // three GEMMs of the recurrence, and carries no accumulated rounding)
        t0 = sector_begin();
        const double* Cx = A;
        const double* Cw = A + N;
        const double* Cp = A + 2 * N;
        upd(W, Cw, nb, 1.0, 0.0, Tm);                          // Tm = P'
        if (p_valid) upd(P, Cp, nb, 1.0, 1.0, Tm);
        QG_HIP(hipMemcpyAsync(P, Tm, blk, hipMemcpyDeviceToDevice, s));
        upd(X, Cx, nb, 1.0, 1.0, P);                           // old P block = X'
        std::swap(X, P);                                       // X = X', P = old X block
        std::swap(P, Tm);                                      // P = P', Tm = free
        hipLaunchKernelGGL(diag_copy_kernel, dim3((N + 255) / 256), dim3(256), 0, s, st.D.p, st.eig.p, N);
        // keep X normalised against rounding
        hipLaunchKernelGGL((coldot_kernel<T>), dim3(N), dim3(256), 0, s, X, X, K, K, st.norm_a.p);
        hipLaunchKernelGGL((colscale_kernel<T>), dim3(grid_for(K) / 4 + 1, N), dim3(256), 0, s, X, (T*)nullptr,
                           st.norm_a.p, K, K);
        sync();
        stats.t_update += sector_end(t0, 16, 1);
        hpsi(X, AX);
        p_valid = true;
        residual(maxres);
        stats.max_residual = maxres;
        if (o.verbose >= 2) std::fprintf(stderr, "lobpcg(ortho) it %d: e0 %.10f max|r| %.3e\n", iter, eig_host[0], maxres);
        if (switch_stall > 0) {
            if (maxres < 0.7 * floor_res) {
                floor_res = maxres;
                floor_iter = iter;
            } else if (iter - floor_iter >= switch_stall) {
                if (o.verbose) std::fprintf(stderr, "lobpcg: residual floor %.3e at iteration %d\n", maxres, iter);
                ++iter;
                break;
            }
        }
        if (totres < best_res) {
            best_res = totres;
            best_iter = iter;
        } else if (learned && o.stall_window > 0 && iter - best_iter >= o.stall_window) {
            learned = false;
            stats.fallback_iter = iter_offset + iter + 1;
            if (o.verbose) std::fprintf(stderr, "lobpcg: learned preconditioner stalled, kinetic from iteration %d\n",
                                        stats.fallback_iter);
        }
    }
    stats.iterations = iter;
    stats.max_residual = maxres;
    // hand back: X (Ritz vectors) normalised with dVol, P in the grid normalisation
    const double isq = 1.0 / sq;
    if constexpr (sizeof(T) == 8) {
        // the blocks were swapped by pointer: move P out of the caller's X array first
        if (P == reinterpret_cast<T*>(X64) && X != reinterpret_cast<T*>(X64)) {
            QG_HIP(hipMemcpyAsync(Tm, P, blk, hipMemcpyDeviceToDevice, s));
            std::swap(P, Tm);
        }
        if (X != reinterpret_cast<T*>(X64)) QG_HIP(hipMemcpyAsync(X64, X, blk, hipMemcpyDeviceToDevice, s));
        if (P != reinterpret_cast<T*>(P64)) QG_HIP(hipMemcpyAsync(P64, P, blk, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL((scale_kernel<double>), dim3(gsz), dim3(256), 0, s, X64, KN, isq);
        if (p_valid) hipLaunchKernelGGL((scale_kernel<double>), dim3(gsz), dim3(256), 0, s, P64, KN, isq);
        else QG_HIP(hipMemsetAsync(P64, 0, blk, s));
    } else {
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(gsz), dim3(256), 0, s, (const float*)X, X64, KN, isq);
        if (p_valid)
            hipLaunchKernelGGL((convert_kernel<float, double>), dim3(gsz), dim3(256), 0, s, (const float*)P, P64, KN,
                               isq);
        else QG_HIP(hipMemsetAsync(P64, 0, blk, s));
    }
    sync();
}

}  // namespace qg

using namespace qg;

// y[k*ld_y + i] = x[k*ld_x + i] (convert between precisions, any leading dimensions)
template <class Ts, class Td>
__global__ void strided_convert_kernel(const Ts* x, int64_t ld_x, Td* y, int64_t ld_y, int64_t ng) {
    const int k = blockIdx.y;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ng; i += (int64_t)gridDim.x * blockDim.x)
        y[(int64_t)k * ld_y + i] = (Td)x[(int64_t)k * ld_x + i];
}

extern "C" {

int qg_gram(int dtype, const void* A, const void* B, int64_t K, int M, int N, int64_t lda, int64_t ldb, double* C,
            int ldc, void* stream) {
    return guarded([&] {
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        QG_REQUIRE(K >= 1 && K < (1LL << 31) && M >= 0 && N >= 0 && lda >= K && ldb >= K && ldc >= M, "bad sizes");
        if (M == 0 || N == 0) return;
        QG_REQUIRE(A && B && C, "null pointer");
        hipStream_t s = as_stream(stream);
        rocblas_handle hb = nullptr;
        QG_BLAS(rocblas_create_handle(&hb));
        struct HandleGuard {
            rocblas_handle h;
            ~HandleGuard() { rocblas_destroy_handle(h); }
        } guard{hb};
        QG_BLAS(rocblas_set_stream(hb, s));
        DevBuf<unsigned char> ws;
        if (dtype == QG_F64) {
            ws.alloc(gram_workspace<double>(K, M, N));
            gram_splitk<double>(hb, s, (const double*)A, lda, (const double*)B, ldb, K, M, N, C, ldc,
                                reinterpret_cast<double*>(ws.p), ws.n);
        } else {
            ws.alloc(gram_workspace<float>(K, M, N));
            gram_splitk<float>(hb, s, (const float*)A, lda, (const float*)B, ldb, K, M, N, C, ldc,
                               reinterpret_cast<float*>(ws.p), ws.n);
        }
        // the workspace is freed on return
        QG_HIP(hipStreamSynchronize(s));
    });
}

int qg_gram_schmidt(int dtype, int64_t ng, int nbands, void* X, int64_t ld, double dvol, void* stream) {
    return guarded([&] {
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        QG_REQUIRE(ng >= 1 && ng < (1LL << 31) && nbands >= 0 && ld >= ng && dvol > 0.0, "bad sizes");
        QG_REQUIRE(nbands <= ng, "more bands than grid points");
        if (nbands == 0) return;
        QG_REQUIRE(X, "null block");
        hipStream_t s = as_stream(stream);
        const int N = nbands;
        rocblas_handle hb = nullptr;
        QG_BLAS(rocblas_create_handle(&hb));
        struct HandleGuard {
            rocblas_handle h;
            ~HandleGuard() { rocblas_destroy_handle(h); }
        } guard{hb};
        QG_BLAS(rocblas_set_stream(hb, s));
        // work in FP64 on a contiguous copy unless the block already is one
        DevBuf<double> tmp;
        double* Y = (double*)X;
        const dim3 cg((unsigned)std::min<int64_t>((ng + 255) / 256, 1024), (unsigned)N);
        if (dtype == QG_F32 || ld != ng) {
            tmp.alloc((size_t)ng * N);
            Y = tmp.p;
            if (dtype == QG_F32)
                hipLaunchKernelGGL((strided_convert_kernel<float, double>), cg, dim3(256), 0, s, (const float*)X, ld, Y, ng, ng);
            else
                hipLaunchKernelGGL((strided_convert_kernel<double, double>), cg, dim3(256), 0, s, (const double*)X, ld, Y, ng, ng);
            QG_HIP(hipGetLastError());
        }
        DevBuf<double> G;
        DevBuf<int> info;
        G.alloc((size_t)N * N);
        info.alloc(1);
        const double one = 1.0, zero = 0.0;
        // two Cholesky-QR passes: G = dVol Y^T Y = R^T R, Y <- Y R^-1 (the second pass restores
        // orthogonality to rounding); R has a positive diagonal, so Y is the Gram-Schmidt basis
        for (int pass = 0; pass < 2; ++pass) {
            QG_BLAS(rocblas_dgemm(hb, rocblas_operation_transpose, rocblas_operation_none, N, N, (int)ng, &dvol, Y,
                                  (int)ng, Y, (int)ng, &zero, G.p, N));
            QG_BLAS(rocsolver_dpotrf(hb, rocblas_fill_upper, N, G.p, N, info.p));
            int h_info = 0;
            QG_HIP(hipMemcpyAsync(&h_info, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
            QG_HIP(hipStreamSynchronize(s));
            if (h_info != 0)
                throw Error{QG_ERR_NUMERIC, "Gram-Schmidt: the bands are linearly dependent (Cholesky info " +
                                                std::to_string(h_info) + ")"};
            QG_BLAS(rocblas_dtrsm(hb, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, (int)ng, N, &one, G.p, N, Y, (int)ng));
        }
        if (Y != (double*)X) {
            if (dtype == QG_F32)
                hipLaunchKernelGGL((strided_convert_kernel<double, float>), cg, dim3(256), 0, s, Y, ng, (float*)X, ld, ng);
            else
                hipLaunchKernelGGL((strided_convert_kernel<double, double>), cg, dim3(256), 0, s, Y, ng, (double*)X, ld, ng);
            QG_HIP(hipGetLastError());
        }
        QG_HIP(hipStreamSynchronize(s));
    });
}

void qg_lobpcg_default_opts(qg_lobpcg_opts* o) {
    if (!o) return;
    o->precision = QG_PREC_F64;
    o->max_iter = 100;
    o->tol = 0.0;
    o->refresh_period = 10;
    o->rollback_limit = 1.0e-3;
    o->precond = QG_PRECOND_NONE;
    o->precond_shift = 1.0;
    o->verbose = 0;
    o->switch_tol = 1.0e-3;
    o->stall_window = 5;
    o->op = nullptr;
    o->op_user = nullptr;
    o->basis = QG_BASIS_REFERENCE;
    o->fallback_shift = 0.0;
    o->watch = nullptr;
    o->watch_user = nullptr;
    o->orth = QG_ORTH_CHOLQR;
    o->gram_chunks = 0;
    o->reortho = 0.0;
    o->nconv = 0;
}

int qg_lobpcg_d_multi(qg_ham* h, const qg_lobpcg_opts* opts, int nbands, double* eig, double* X, double* P,
                      int scf_step, qg_lobpcg_stats* stats) {
    return guarded([&] {
        QG_REQUIRE(h && eig && X && P, "null argument");
        QG_REQUIRE(nbands >= 1, "need at least one band");
        QG_REQUIRE((int64_t)3 * nbands <= h->grid.ng(), "3N must not exceed the grid size");
        qg_lobpcg_opts o;
        qg_lobpcg_default_opts(&o);
        if (opts) o = *opts;
        QG_REQUIRE(o.precision == QG_PREC_F64 || o.precision == QG_PREC_F32 || o.precision == QG_PREC_MIXED,
                   "unknown precision mode");
        QG_REQUIRE(o.precond == QG_PRECOND_NONE || o.precond == QG_PRECOND_KINETIC || o.precond == QG_PRECOND_LEARNED,
                   "unknown preconditioner");
        QG_REQUIRE(o.precond != QG_PRECOND_LEARNED || h->cnn, "QG_PRECOND_LEARNED needs a network (qg_hamiltonian_set_cnn)");
        QG_REQUIRE(o.max_iter >= 0 && o.rollback_limit > 0.0, "bad iteration options");
        QG_REQUIRE(o.nconv >= 0, "nconv must not be negative");
        QG_REQUIRE(h->grid.ng() < (1LL << 31), "grid too large for 32-bit BLAS dimensions");
        qg_lobpcg_stats st{};
        st.fallback_iter = -1;
        const bool reference = o.precision == QG_PREC_F64 && o.precond == QG_PRECOND_NONE && o.basis == QG_BASIS_REFERENCE;
        if (reference) {
            run_lobpcg<double>(h, o, nbands, eig, X, P, scf_step, st);
        } else if (o.precision == QG_PREC_F64) {
            run_lobpcg_ortho<double>(h, o, nbands, eig, X, P, scf_step, st);
        } else {
            // FP32 phase down to switch_tol, then FP64 refinement sweeps down to tol
            const bool refine = o.tol > 0.0 && o.tol < o.switch_tol;
            qg_lobpcg_opts o1 = o;
            if (refine) o1.tol = o.switch_tol;
            // the FP32 phase also ends at its rounding floor (residual stalled)
            run_lobpcg_ortho<float>(h, o1, nbands, eig, X, P, scf_step, st, 0,
                                    refine ? 3 * std::max(o.stall_window, 3) : 0);
            if (refine && st.max_residual > o.tol && st.iterations < o.max_iter) {
                qg_lobpcg_opts o2 = o;
                o2.precision = QG_PREC_F64;
                o2.max_iter = o.max_iter - st.iterations;
                qg_lobpcg_stats s2{};
                s2.fallback_iter = st.fallback_iter;
                run_lobpcg_ortho<double>(h, o2, nbands, eig, X, P, 0, s2, st.iterations);
                st.fallback_iter = s2.fallback_iter;
                st.iterations += s2.iterations;
                st.rollbacks += s2.rollbacks;
                st.hpsi_calls += s2.hpsi_calls;
                st.max_residual = s2.max_residual;
                st.t_hpsi += s2.t_hpsi; st.t_gram += s2.t_gram; st.t_update += s2.t_update;
                st.t_eig += s2.t_eig; st.t_other += s2.t_other; st.t_precond += s2.t_precond;
                st.refine_iterations = s2.iterations;
            }
        }
        if (stats) *stats = st;
    });
}

int qg_batch_lobpcg(int ncells, qg_ham* const* hams, const qg_lobpcg_opts* opts, const int* nbands,
                    double* const* eig, double* const* X, double* const* P, int scf_step, qg_lobpcg_stats* stats,
                    int batched) {
    return guarded([&] {
        QG_REQUIRE(ncells >= 1 && hams && nbands && eig && X && P, "bad arguments");
        for (int c = 0; c < ncells; ++c) {
            QG_REQUIRE(hams[c] && eig[c] && X[c] && P[c] && nbands[c] >= 1, "bad cell");
            for (int d = 0; d < c; ++d) QG_REQUIRE(hams[d] != hams[c], "a Hamiltonian appears twice in the batch");
        }
        if (!batched) {
            // one cell after another, each with its own launches
            for (int c = 0; c < ncells; ++c) {
                const int rc = qg_lobpcg_d_multi(hams[c], opts, nbands[c], eig[c], X[c], P[c], scf_step,
                                                 stats ? &stats[c] : nullptr);
                if (rc != QG_OK) throw Error{rc, std::string("cell ") + std::to_string(c) + ": " + qg_last_error()};
            }
            return;
        }
        BatchService svc;
        svc.active = ncells;
        for (int c = 0; c < ncells; ++c) hams[c]->batch = &svc;
        std::vector<std::string> errs(ncells);
        std::vector<int> codes(ncells, QG_OK);
        std::vector<std::thread> th;
        for (int c = 0; c < ncells; ++c)
            th.emplace_back([&, c] {
                qg_lobpcg_stats st{};
                codes[c] = qg_lobpcg_d_multi(hams[c], opts, nbands[c], eig[c], X[c], P[c], scf_step, &st);
                if (codes[c] != QG_OK) errs[c] = qg_last_error();
                if (stats) stats[c] = st;
                hipStream_t s = hams[c]->lobpcg ? hams[c]->lobpcg->stream : nullptr;
                batch_service_leave(&svc, s);
            });
        for (auto& t : th) t.join();
        for (int c = 0; c < ncells; ++c) hams[c]->batch = nullptr;
        for (int c = 0; c < ncells; ++c)
            if (codes[c] != QG_OK) throw Error{codes[c], std::string("cell ") + std::to_string(c) + ": " + errs[c]};
    });
}

}  // extern "C"
