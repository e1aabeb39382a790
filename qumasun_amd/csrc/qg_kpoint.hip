// qg_kpoint.hip — the k-point (complex Bloch) Hamiltonian and its LOBPCG on one GPU.
//
// The cell-periodic part u of psi_k = exp(i k.r) u, band-major interleaved complex on the
// device (re, im adjacent; ld in complex elements), is acted on by
//   H_k u = coef [L8 u - |k|^2 u + 2 i k . D8 u] + V u + sum_j E_j conj(P_j) (dV P_j^T u),
// coef = -1/2, with P_j = beta_j(x_u - R) exp(i k . x_u) on the projector's image blocks.
//
// Reference semantics (QUMASUN source/):
//   stencil   LaplasianKpoint8th_halo_f on re and im         GridDifferenceKpoint8th.h:17-148
//             (c0 carries -coef |k|^2; the first-derivative term 2 coef k . D8 moves re -> im with
//             sign +1 and im -> re with sign -1)
//   operator  mHamiltonianMatrix_ddm at a k point            qumasun_kpoint.matrix.h:21-110
//   projectors mCreateProjectorRadialOnBlock: beta Ylm (cos, sin)(k . x_u), x_u = h (i - shift)
//                                                          PseudoPotOperator_kpoint_v2.h:287-420
//   project   mInnerNonlocalPsi_v2: c = dV sum P u (no conjugate)       :836-930
//   apply     ProjectionPP: Hp += E conj(P) c                           :2581-2690
//             (gathered per grid point over the projector entries that touch it: no atomics)
//   solver    Eigen_z_multi_mpi (ZGEMM Grams, ZHEGVX)       lobpcg_z_multi_mpi.h:988-1600
//             restated on an orthonormal basis: the Rayleigh-Ritz problem of [X W P] is a
//             standard Hermitian one (zheevd), so no B-metric solve is needed
#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocfft/rocfft.h>
#include <rocsolver/rocsolver.h>
#include <map>
#include <memory>
#include <mutex>
#include "qg_internal.hpp"

#define QG_KBLAS(call)                                                                       \
    do {                                                                                     \
        rocblas_status st_ = (call);                                                         \
        if (st_ != rocblas_status_success)                                                   \
            throw ::qg::Error{QG_ERR_LIB, std::string(#call) + ": " + rocblas_status_to_string(st_)}; \
    } while (0)

#define QG_KFFT(call)                                                                          \
    do {                                                                                       \
        rocfft_status st_ = (call);                                                            \
        if (st_ != rocfft_status_success)                                                      \
            throw ::qg::Error{QG_ERR_LIB, std::string(#call) + " failed: " + std::to_string((int)st_)}; \
    } while (0)

// complex-to-complex FFT plans of one batch size for the k-point kinetic preconditioner
struct KFftPlan {
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    qg::DevBuf<unsigned char> work;
    ~KFftPlan() {
        if (fwd) rocfft_plan_destroy(fwd);
        if (inv) rocfft_plan_destroy(inv);
        if (info) rocfft_execution_info_destroy(info);
    }
};

struct qg_kpoint {
    qg_ham* h = nullptr;   // not owned
    double k[3] = {0.0, 0.0, 0.0};
    int geo_version = -1;  // projector geometry the phase tables were built for
    qg::DevBuf<double2> phase64;
    qg::DevBuf<float2> phase32;
    qg::DevBuf<double2> coef;   // projector coefficients [nb][ncoef]
    // the projector table entries of every grid point (CSR over grid points): entries of point
    // g are pt_ent[pt_off[g] .. pt_off[g+1]), each a table point and its atom's first coefficient
    qg::DevBuf<int> pt_off, pt_ent, pt_cb;
    qg::DevBuf<unsigned char> ec;   // E c per band, T2, rows of ncoef + 8 (zero padded)
    // kinetic preconditioner: plans per batch size, spectrum buffer, per-axis symbol parts
    std::map<int, std::unique_ptr<KFftPlan>> fft;
    qg::DevBuf<double2> spec;
    qg::DevBuf<double> tx, ty, tz;
    double tmin = 0.0;   // smallest value of the symbol over all modes
    // LOBPCG state
    rocblas_handle blas = nullptr;
    ~qg_kpoint() {
        if (blas) (void)rocblas_destroy_handle(blas);
    }
};

namespace qg {
namespace {

constexpr int kThreads = 256;
constexpr int kTX = 32, kTY = 8;
constexpr int kHX = kTX + 8, kHY = kTY + 8;   // the tile with its 4-point halo
constexpr int kHalo = kHX * kHY;
constexpr int kLoadsPT = (kHalo + kThreads - 1) / kThreads;
constexpr int kXcds = 8;

template <class T> struct Cx;
template <> struct Cx<float> { using t = float2; };
template <> struct Cx<double> { using t = double2; };

__device__ __forceinline__ int wrap_any(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// stencil weights: centre, and per axis the second-difference (a) and first-difference (b)
// weights of offsets 1..4
template <class T>
struct KStencil {
    T c0;
    T a[3][4];
    T b[3][4];
};

// out = coef [L8 - |k|^2 + 2 i k . D8] u + V u for one band per workgroup column tile.  The
// tile (kTX x kTY points, one per thread) walks the planes z' = -4 .. nz+3 (periodic images at
// both ends): each plane of the tile and its halo is read from global memory once, two planes
// ahead of its use, into one of two LDS buffers (one barrier per plane).  The x and y terms of
// plane z' are formed from LDS and queued for four planes; plane z'-4 is finished from that
// queue, the register queue of the thread's nine column values (z terms) and V.  The flat
// workgroup index is remapped so that each XCD works on a contiguous range of (band, tile)
// pairs.
template <class T>
__global__ __launch_bounds__(kThreads) void kstencil_kernel(const typename Cx<T>::t* __restrict__ psi, int64_t ld,
                                                             const T* __restrict__ V, int nx, int ny, int nz,
                                                             int tiles_x, int ntiles, int nwork, int per_xcd,
                                                             KStencil<T> ks, typename Cx<T>::t* __restrict__ out,
                                                             int64_t ld_out) {
    using C = typename Cx<T>::t;
    // the tile with its halo, padded to one slot per load so that loads and stores carry no
    // conditions (the compiler then keeps the next plane's loads in flight across them); real
    // and imaginary parts in separate planes and registers (no vector types held in registers)
    __shared__ T bre[2][kLoadsPT * kThreads];
    __shared__ T bim[2][kLoadsPT * kThreads];
    const int lin = blockIdx.x;
    const int work = (lin % kXcds) * per_xcd + lin / kXcds;
    if (work >= nwork) return;   // padding of the remap: the whole workgroup leaves
    const int band = work / ntiles;
    const int tile = work - band * ntiles;
    const int x0 = (tile % tiles_x) * kTX, y0 = (tile / tiles_x) * kTY;
    const T* p = reinterpret_cast<const T*>(psi + (int64_t)band * ld);
    const int64_t sxy = (int64_t)nx * ny;
    int64_t src[kLoadsPT];   // offsets in T elements (2 per point)
#pragma unroll
    for (int k = 0; k < kLoadsPT; ++k) {
        const int e = threadIdx.x + k * kThreads;
        const int hy = e / kHX, hx = e - hy * kHX;
        src[k] = 2 * (e < kHalo ? (int64_t)wrap_any(x0 - 4 + hx, nx) + (int64_t)nx * wrap_any(y0 - 4 + hy, ny) : 0);
    }
    const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
    const bool active = x0 + tx < nx && y0 + ty < ny;
    const int c = (ty + 4) * kHX + tx + 4;
    const int64_t pt0 = (x0 + tx) + (int64_t)nx * (y0 + ty);
    const int64_t ob = (int64_t)band * ld_out + pt0;
    const int zend = nz + 4 + (nz & 1);   // even step count; the pad step writes nothing
    // the weights stay in the kernel argument (uniform: scalar registers)
    // column values of planes z'-8 .. z' and the x and y terms of planes z'-4 .. z'
    T qr[9], qi[9], pr[5], pi[5];
#pragma unroll
    for (int k = 0; k < 9; ++k) qr[k] = qi[k] = T(0);
#pragma unroll
    for (int k = 0; k < 5; ++k) pr[k] = pi[k] = T(0);
    T ar_[kLoadsPT], ai_[kLoadsPT], br_[kLoadsPT], bi_[kLoadsPT];   // two planes in flight

#define QG_KFETCH(z, vr, vi)                                     \
    {                                                            \
        const int64_t zo = 2 * sxy * wrap_any((z), nz);          \
        _Pragma("unroll") for (int k = 0; k < kLoadsPT; ++k) {   \
            vr[k] = p[zo + src[k]];                              \
            vi[k] = p[zo + src[k] + 1];                          \
        }                                                        \
    }
    // one axis: sum_m a_m (u+ + u-) + i b_m (u+ - u-)
    auto axis = [&](T& sr, T& si, T upx, T upy, T dnx, T dny, T am, T bm) __attribute__((always_inline)) {
        sr += am * (upx + dnx) - bm * (upy - dny);
        si += am * (upy + dny) + bm * (upx - dnx);
    };
    // plane z' arrives in (vr, vi): stage it, fetch plane z'+2, difference what is ready
#define QG_KSTEP(zq, vr, vi)                                                                           \
    {                                                                                                  \
        const int z = (zq);                                                                            \
        T* br = bre[z & 1];                                                                            \
        T* bi = bim[z & 1];                                                                            \
        _Pragma("unroll") for (int k = 0; k < kLoadsPT; ++k) {                                         \
            br[threadIdx.x + k * kThreads] = vr[k];                                                    \
            bi[threadIdx.x + k * kThreads] = vi[k];                                                    \
        }                                                                                              \
        QG_KFETCH(z + 2, vr, vi) /* unconditional (periodic, in bounds): the waits stay counted */     \
        __syncthreads();                                                                               \
        _Pragma("unroll") for (int k = 0; k < 8; ++k) {                                                \
            qr[k] = qr[k + 1];                                                                         \
            qi[k] = qi[k + 1];                                                                         \
        }                                                                                              \
        qr[8] = br[c];                                                                                 \
        qi[8] = bi[c];                                                                                 \
        _Pragma("unroll") for (int k = 0; k < 4; ++k) {                                                \
            pr[k] = pr[k + 1];                                                                         \
            pi[k] = pi[k + 1];                                                                         \
        }                                                                                              \
        T sr = T(0), si = T(0);                                                                        \
        if (z >= 0 && z < nz) {                                                                        \
            _Pragma("unroll") for (int m = 1; m <= 4; ++m) {                                           \
                axis(sr, si, br[c + m], bi[c + m], br[c - m], bi[c - m], ks.a[0][m - 1], ks.b[0][m - 1]); \
                axis(sr, si, br[c + m * kHX], bi[c + m * kHX], br[c - m * kHX], bi[c - m * kHX],        \
                     ks.a[1][m - 1], ks.b[1][m - 1]);                                                  \
            }                                                                                          \
        }                                                                                              \
        pr[4] = sr;                                                                                    \
        pi[4] = si;                                                                                    \
        if (active && z >= 4 && z - 4 < nz) {                                                          \
            const int zz = z - 4;                                                                      \
            T orr = pr[0], oii = pi[0];                                                                \
            _Pragma("unroll") for (int m = 1; m <= 4; ++m)                                             \
                axis(orr, oii, qr[4 + m], qi[4 + m], qr[4 - m], qi[4 - m], ks.a[2][m - 1], ks.b[2][m - 1]); \
            const T d = ks.c0 + V[pt0 + sxy * zz];                                                     \
            orr += d * qr[4];                                                                          \
            oii += d * qi[4];                                                                          \
            out[ob + sxy * zz] = C{orr, oii};                                                          \
        }                                                                                              \
    }
    QG_KFETCH(-4, ar_, ai_)
    QG_KFETCH(-3, br_, bi_)
    for (int z0 = -4; z0 < zend; z0 += 2) {
        QG_KSTEP(z0, ar_, ai_)
        QG_KSTEP(z0 + 1, br_, bi_)
    }
#undef QG_KSTEP
#undef QG_KFETCH
}

// k-point projection and application over the per-point projector tables: one workgroup per
// (atom, chunk of NBA bands) streaming the atom's table points.
struct KProjArgs {
    int natoms;
    int nbands;
    int64_t ld;
    double dv;
    const int* atom_img_off;
    const int* atom_coef_off;
    const int64_t* img_tbl_off;
    const int* gidx;
    const void* beta;      // [point][NJS] beta Ylm in T
    const void* phase;     // [point] (cos, sin)(k . x_u) in T
    const double* energy;  // per coefficient
    double2* coef;         // [band][ncoef]
    int ncoef;
};

constexpr int kProjBlock = 256;

__device__ __forceinline__ int xcd_item(int i, int n) {
    // consecutive items (neighbouring atoms of one band chunk) on one XCD
    const int per = (n + kXcds - 1) / kXcds;
    const int j = (i % kXcds) * per + i / kXcds;
    return j;
}

// c[b][j] = dV sum_i beta_j(i) phase(i) u_b(g(i))   (no conjugate)
template <class T, int NBA, int NJS>
__global__ __launch_bounds__(kProjBlock) void kproject_kernel(KProjArgs a, const typename Cx<T>::t* __restrict__ psi,
                                                               int nitems) {
    using C = typename Cx<T>::t;
    __shared__ double red[kProjBlock / 64][NJS * NBA * 2];
    const int item = xcd_item(blockIdx.x, gridDim.x);
    if (item >= nitems) return;   // padding of the remap: the whole workgroup leaves
    const int atom = item % a.natoms;
    const int chunk = item / a.natoms;
    const int b0 = chunk * NBA;
    const int tid = threadIdx.x;
    const T* beta = reinterpret_cast<const T*>(a.beta);
    const C* ph = reinterpret_cast<const C*>(a.phase);
    const C* src[NBA];
#pragma unroll
    for (int b = 0; b < NBA; ++b) src[b] = psi + (int64_t)min(b0 + b, a.nbands - 1) * a.ld;
    T ar[NJS][NBA], ai[NJS][NBA];
#pragma unroll
    for (int j = 0; j < NJS; ++j)
#pragma unroll
        for (int b = 0; b < NBA; ++b) ar[j][b] = ai[j][b] = T(0);
    const int64_t t0 = a.img_tbl_off[a.atom_img_off[atom]];
    const int64_t t1 = a.img_tbl_off[a.atom_img_off[atom + 1]];
    for (int64_t i = t0 + tid; i < t1; i += kProjBlock) {
        const int g = a.gidx[i];
        const C e = ph[i];
        T bj[NJS];
#pragma unroll
        for (int j = 0; j < NJS; ++j) bj[j] = beta[i * NJS + j];
#pragma unroll
        for (int b = 0; b < NBA; ++b) {
            const C u = src[b][g];
            const T wr = e.x * u.x - e.y * u.y, wi = e.x * u.y + e.y * u.x;
#pragma unroll
            for (int j = 0; j < NJS; ++j) {
                ar[j][b] += bj[j] * wr;
                ai[j][b] += bj[j] * wi;
            }
        }
    }
    const int wave = tid >> 6;
#pragma unroll
    for (int j = 0; j < NJS; ++j)
#pragma unroll
        for (int b = 0; b < NBA; ++b) {
            double vr = (double)ar[j][b], vi = (double)ai[j][b];
#pragma unroll
            for (int s = 32; s >= 1; s >>= 1) {
                vr += __shfl_xor(vr, s, 64);
                vi += __shfl_xor(vi, s, 64);
            }
            if ((tid & 63) == 0) {
                red[wave][(j * NBA + b) * 2] = vr;
                red[wave][(j * NBA + b) * 2 + 1] = vi;
            }
        }
    __syncthreads();
    if (tid < NJS * NBA) {
        const int j = tid / NBA, b = tid % NBA;
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int w = 0; w < kProjBlock / 64; ++w) {
            sr += red[w][tid * 2];
            si += red[w][tid * 2 + 1];
        }
        const int cbase = a.atom_coef_off[atom];
        const int count = a.atom_coef_off[atom + 1] - cbase;
        if (b0 + b < a.nbands && j < count)
            a.coef[(int64_t)(b0 + b) * a.ncoef + cbase + j] = double2{sr * a.dv, si * a.dv};
    }
}

// ec[b][j] = E_j c[b][j] in the states' precision, rows of ld_ec = ncoef + 8 with zero padding
template <class T>
__global__ void kscale_coef_kernel(const double2* __restrict__ c, const double* __restrict__ energy, int ncoef,
                                   int ld_ec, int nbands, typename Cx<T>::t* __restrict__ ec) {
    const int64_t n = (int64_t)nbands * ld_ec;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(t / ld_ec), j = (int)(t - (int64_t)b * ld_ec);
        typename Cx<T>::t v{T(0), T(0)};
        if (j < ncoef) {
            const double2 cc = c[(int64_t)b * ncoef + j];
            v.x = (T)(energy[j] * cc.x);
            v.y = (T)(energy[j] * cc.y);
        }
        ec[t] = v;
    }
}

// out_b(g) += sum over the table entries of g of conj(phase) sum_j beta_j E_j c[b][cb + j]: one
// thread per grid point and chunk of NB bands, one read-modify-write of the output per point and
// band (the projector points of all atoms that touch g are gathered, so no two threads write one
// value).  Slots j past an atom's coefficients read the next atom's (or the padding's) finite
// values against beta_j = 0.
template <class T, int NB, int NJS>
__global__ __launch_bounds__(256) void kgather_kernel(const int* __restrict__ pt_off, const int* __restrict__ pt_ent,
                                                       const int* __restrict__ pt_cb, const T* __restrict__ beta,
                                                       const typename Cx<T>::t* __restrict__ phase,
                                                       const typename Cx<T>::t* __restrict__ ec, int ld_ec, int nbands,
                                                       int64_t ng, typename Cx<T>::t* __restrict__ out,
                                                       int64_t ld_out) {
    using C = typename Cx<T>::t;
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int e0 = pt_off[g], e1 = pt_off[g + 1];
    if (e0 == e1) return;
    const int b0 = blockIdx.y * NB;
    const int nb = min(NB, nbands - b0);
    T ar[NB], ai[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) ar[b] = ai[b] = T(0);
    for (int e = e0; e < e1; ++e) {
        const int t = pt_ent[e];
        const int cb = pt_cb[e];
        const C ph = phase[t];
        T bj[NJS];
#pragma unroll
        for (int j = 0; j < NJS; ++j) bj[j] = beta[(int64_t)t * NJS + j];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (b >= nb) break;
            const C* row = ec + (int64_t)(b0 + b) * ld_ec + cb;
            T vr = T(0), vi = T(0);
#pragma unroll
            for (int j = 0; j < NJS; ++j) {
                const C w = row[j];
                vr += bj[j] * w.x;
                vi += bj[j] * w.y;
            }
            ar[b] += ph.x * vr + ph.y * vi;
            ai[b] += ph.x * vi - ph.y * vr;
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (b >= nb) break;
        C* o = out + (int64_t)(b0 + b) * ld_out + g;
        C v = *o;
        v.x += ar[b];
        v.y += ai[b];
        *o = v;
    }
}

// bands per projector workgroup: each table point's beta, phase and grid index are read once
// per chunk, so FP32 takes 8 (FP64 keeps 4 to stay within registers)
template <class T>
constexpr int kNBA = sizeof(T) == 4 ? 8 : 4;

// phase tables of the k point on the projector points: (cos, sin)(k . h (i - shift))
void build_phases(qg_kpoint* kp) {
    qg_nonlocal* nl = kp->h->nl;
    if (!nl || kp->geo_version == nl->geo_version) return;
    const NlGeometry& geo = nl->geo;
    const qg_grid& g = nl->gcopy;
    const int64_t np = geo.img_tbl_off.empty() ? 0 : geo.img_tbl_off.back();
    std::vector<double2> p64((size_t)np);
    std::vector<float2> p32((size_t)np);
    for (size_t im = 0; im + 1 < geo.img_tbl_off.size(); ++im) {
        const DevImage& img = geo.images[im];
        for (int64_t t = geo.img_tbl_off[im]; t < geo.img_tbl_off[im + 1]; ++t) {
            const int gi = geo.gidx[t];
            const int ix = gi % g.n[0];
            const int iy = (gi / g.n[0]) % g.n[1];
            const int iz = gi / (g.n[0] * g.n[1]);
            const double th = kp->k[0] * g.h[0] * (double)(ix - img.shift[0]) +
                              kp->k[1] * g.h[1] * (double)(iy - img.shift[1]) +
                              kp->k[2] * g.h[2] * (double)(iz - img.shift[2]);
            p64[t] = double2{std::cos(th), std::sin(th)};
            p32[t] = float2{(float)p64[t].x, (float)p64[t].y};
        }
    }
    if (np) {
        kp->phase64.upload(p64.data(), p64.size());
        kp->phase32.upload(p32.data(), p32.size());
    }
    // the entries of every grid point (CSR), with their atom's first coefficient
    const int64_t ngp = g.ng();
    std::vector<int> off((size_t)ngp + 1, 0), ent((size_t)np), cbv((size_t)np);
    for (int64_t t = 0; t < np; ++t) ++off[(size_t)geo.gidx[t] + 1];
    for (int64_t i = 0; i < ngp; ++i) off[(size_t)i + 1] += off[(size_t)i];
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (size_t im = 0; im + 1 < geo.img_tbl_off.size(); ++im) {
        const int cb = geo.atom_coef_off[geo.images[im].atom];
        for (int64_t t = geo.img_tbl_off[im]; t < geo.img_tbl_off[im + 1]; ++t) {
            const int pos = fill[geo.gidx[t]]++;
            ent[pos] = (int)t;
            cbv[pos] = cb;
        }
    }
    kp->pt_off.upload(off.data(), off.size());
    if (np) {
        kp->pt_ent.upload(ent.data(), ent.size());
        kp->pt_cb.upload(cbv.data(), cbv.size());
    }
    kp->geo_version = nl->geo_version;
}

template <class T>
void launch_nl(qg_kpoint* kp, int nbands, const typename Cx<T>::t* psi, int64_t ld, typename Cx<T>::t* out,
               int64_t ld_out, hipStream_t s) {
    qg_nonlocal* nl = kp->h->nl;
    const NlGeometry& geo = nl->geo;
    QG_REQUIRE(geo.max_coef_per_atom <= 8, "k-point projectors: at most 8 coefficients per atom");
    kp->coef.alloc((size_t)nbands * geo.ncoef);
    KProjArgs a;
    a.natoms = geo.natoms;
    a.nbands = nbands;
    a.ld = ld;
    a.dv = nl->gcopy.dvol();
    a.atom_img_off = nl->d_atom_img_off.p;
    a.atom_coef_off = nl->d_atom_coef_off.p;
    a.img_tbl_off = nl->d_img_tbl_off.p;
    a.gidx = nl->d_gidx.p;
    a.beta = sizeof(T) == 4 ? (const void*)nl->d_beta32.p : (const void*)nl->d_beta64.p;
    a.phase = sizeof(T) == 4 ? (const void*)kp->phase32.p : (const void*)kp->phase64.p;
    a.energy = nl->d_coef_energy.p;
    a.coef = kp->coef.p;
    a.ncoef = geo.ncoef;
    const int64_t chunks = (nbands + kNBA<T> - 1) / kNBA<T>;
    const int64_t items = (int64_t)geo.natoms * chunks;
    QG_REQUIRE(items < (int64_t(1) << 31) - kXcds, "too many projector items in one launch");
    auto padded = [](int64_t n) { return (unsigned)(((n + kXcds - 1) / kXcds) * kXcds); };
    if (geo.njs == 4)
        hipLaunchKernelGGL((kproject_kernel<T, kNBA<T>, 4>), dim3(padded(items)), dim3(kProjBlock), 0, s, a, psi,
                           (int)items);
    else
        hipLaunchKernelGGL((kproject_kernel<T, kNBA<T>, 8>), dim3(padded(items)), dim3(kProjBlock), 0, s, a, psi,
                           (int)items);
    QG_HIP(hipGetLastError());
    // the application: E c per band (padded rows), then one gather launch
    const int ld_ec = geo.ncoef + 8;
    using C = typename Cx<T>::t;
    kp->ec.alloc((size_t)nbands * ld_ec * sizeof(C));
    C* ec = reinterpret_cast<C*>(kp->ec.p);
    hipLaunchKernelGGL(kscale_coef_kernel<T>, dim3((unsigned)std::min<int64_t>(((int64_t)nbands * ld_ec + 255) / 256, 4096)),
                       dim3(256), 0, s, (const double2*)kp->coef.p, nl->d_coef_energy.p, geo.ncoef, ld_ec, nbands, ec);
    const int64_t ng = nl->gcopy.ng();
    constexpr int NB = sizeof(T) == 4 ? 8 : 4;
    const dim3 grid((unsigned)((ng + 255) / 256), (unsigned)((nbands + NB - 1) / NB));
    QG_REQUIRE(grid.y <= 65535, "too many bands in one k-point application");
    const T* beta = sizeof(T) == 4 ? (const T*)nl->d_beta32.p : (const T*)nl->d_beta64.p;
    const C* ph = sizeof(T) == 4 ? (const C*)kp->phase32.p : (const C*)kp->phase64.p;
    if (geo.njs == 4)
        hipLaunchKernelGGL((kgather_kernel<T, NB, 4>), grid, dim3(256), 0, s, kp->pt_off.p, kp->pt_ent.p, kp->pt_cb.p,
                           beta, ph, ec, ld_ec, nbands, ng, out, ld_out);
    else
        hipLaunchKernelGGL((kgather_kernel<T, NB, 8>), grid, dim3(256), 0, s, kp->pt_off.p, kp->pt_ent.p, kp->pt_cb.p,
                           beta, ph, ec, ld_ec, nbands, ng, out, ld_out);
    QG_HIP(hipGetLastError());
}

template <class T>
void kpoint_apply(qg_kpoint* kp, int nbands, const void* psi, int64_t ld, void* out, int64_t ld_out, hipStream_t s) {
    qg_ham* h = kp->h;
    const qg_grid& g = h->grid;
    const double coef = -0.5;
    KStencil<T> ks;
    const double k2 = kp->k[0] * kp->k[0] + kp->k[1] * kp->k[1] + kp->k[2] * kp->k[2];
    ks.c0 = (T)(-14350.0 / 5040.0 * coef * (1.0 / (g.h[0] * g.h[0]) + 1.0 / (g.h[1] * g.h[1]) + 1.0 / (g.h[2] * g.h[2])) -
                coef * k2);
    const double c2[4] = {8064.0, -1008.0, 128.0, -9.0};
    const double c1[4] = {672.0, -168.0, 32.0, -3.0};
    for (int d = 0; d < 3; ++d)
        for (int m = 0; m < 4; ++m) {
            ks.a[d][m] = (T)(c2[m] * coef / (5040.0 * g.h[d] * g.h[d]));
            ks.b[d][m] = (T)(2.0 * coef * kp->k[d] * c1[m] / (840.0 * g.h[d]));
        }
    const int tiles_x = (g.n[0] + kTX - 1) / kTX, tiles_y = (g.n[1] + kTY - 1) / kTY;
    const int64_t ntiles = (int64_t)tiles_x * tiles_y;
    const int64_t nwork = ntiles * nbands;
    QG_REQUIRE(nwork < (int64_t(1) << 31) - kXcds, "too many stencil tiles in one launch");
    const int64_t per_xcd = (nwork + kXcds - 1) / kXcds;
    using C = typename Cx<T>::t;
    const T* V = sizeof(T) == 4 ? (const T*)h->V32.p : (const T*)h->V64.p;
    hipLaunchKernelGGL(kstencil_kernel<T>, dim3((unsigned)(per_xcd * kXcds)), dim3(kThreads), 0, s, (const C*)psi,
                       ld, V, g.n[0], g.n[1], g.n[2], tiles_x, (int)ntiles, (int)nwork, (int)per_xcd, ks, (C*)out,
                       ld_out);
    QG_HIP(hipGetLastError());
    if (h->nl && h->nl->geo.ncoef > 0) {
        build_phases(kp);
        launch_nl<T>(kp, nbands, (const C*)psi, ld, (C*)out, ld_out, s);
    }
}

void check_kblock(const qg_kpoint* kp, int dtype, int nbands, const void* psi, int64_t ld, const void* out,
                  int64_t ld_out) {
    QG_REQUIRE(kp && kp->h, "null k-point handle");
    QG_REQUIRE(dtype == QG_Z128 || dtype == QG_C64, "k-point blocks are complex: QG_Z128 or QG_C64");
    QG_REQUIRE(nbands >= 0, "negative band count");
    const int64_t ng = kp->h->grid.ng();
    QG_REQUIRE(ld >= ng && ld_out >= ng, "leading dimension smaller than the grid");
    if (nbands > 0) {
        QG_REQUIRE(psi && out, "null pointer");
        QG_REQUIRE(psi != out, "in-place application is not supported (psi must differ from out)");
    }
}

// ----------------------------------------------------------- kinetic preconditioner
// K = (T_k + s)^-1 with T_k = -1/2 [L8 - |k|^2 + 2 i k . D8], the k-point kinetic operator of
// the stencil kernel.  On the periodic grid it is diagonal in the discrete Fourier basis: for the
// mode exp(i theta_d j_d) per axis, T_k = sum_d t_d(theta_d) with
//   t_d = -1/2 h^-2 [w0 + 2 sum_m w_m cos(m theta)] + 1/2 k_d^2 + k_d (2 / (840 h)) sum_m c_m sin(m theta),
// (w0..w4) = (-205/72, 8/5, -1/5, 8/315, -1/560), (c1..c4) = (672, -168, 32, -3), so K is one
// complex FFT, a pointwise scale by 1 / (T_k + s) / Ng and the inverse FFT, batched over bands.
static std::vector<double> kinetic_symbol_k(int n, double h, double k) {
    const double w[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
    const double c[4] = {672.0, -168.0, 32.0, -3.0};
    std::vector<double> t(n);
    for (int q = 0; q < n; ++q) {
        const double th = 2.0 * M_PI * q / n;
        double s2 = w[0], s1 = 0.0;
        for (int m = 1; m <= 4; ++m) {
            s2 += 2.0 * w[m] * std::cos(m * th);
            s1 += c[m - 1] * std::sin(m * th);
        }
        t[q] = -0.5 * s2 / (h * h) + 0.5 * k * k + k * 2.0 * s1 / (840.0 * h);
    }
    return t;
}

__global__ void kinetic_scale_c2c_kernel(double2* spec, int nx, int ny, int nz, int nb, const double* tx,
                                         const double* ty, const double* tz, double shift, double inv_ng) {
    const int64_t per = (int64_t)nx * ny * nz;
    const int64_t total = per * nb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i % per;
        const int kx = (int)(j % nx);
        const int ky = (int)((j / nx) % ny);
        const int kz = (int)(j / ((int64_t)nx * ny));
        const double f = inv_ng / (tx[kx] + ty[ky] + tz[kz] + shift);
        spec[i].x *= f;
        spec[i].y *= f;
    }
}

void ensure_rocfft() {
    static std::once_flag once;
    std::call_once(once, [] { rocfft_setup(); });
}

// R <- K R for nb complex FP64 bands (ld = Ng)
void kinetic_precond(qg_kpoint* kp, int nb, double2* R, double shift, hipStream_t s) {
    ensure_rocfft();
    const qg_grid& g = kp->h->grid;
    const int64_t ng = g.ng();
    if (!kp->tx.p) {
        auto a = kinetic_symbol_k(g.n[0], g.h[0], kp->k[0]);
        auto b = kinetic_symbol_k(g.n[1], g.h[1], kp->k[1]);
        auto c = kinetic_symbol_k(g.n[2], g.h[2], kp->k[2]);
        kp->tmin = *std::min_element(a.begin(), a.end()) + *std::min_element(b.begin(), b.end()) +
                   *std::min_element(c.begin(), c.end());
        kp->tx.upload(a.data(), a.size());
        kp->ty.upload(b.data(), b.size());
        kp->tz.upload(c.data(), c.size());
    }
    QG_REQUIRE(kp->tmin + shift > 0.0, "k-point preconditioner: T_k + shift is not positive on every mode");
    // bands per FFT batch: about 256 MB of spectrum
    const int batch = (int)std::max<int64_t>(1, std::min<int64_t>(nb, (256ll << 20) / (ng * 16)));
    kp->spec.alloc((size_t)batch * ng);
    for (int b0 = 0; b0 < nb; b0 += batch) {
        const int nc = std::min(batch, nb - b0);
        auto& slot = kp->fft[nc];
        if (!slot) {
            slot.reset(new KFftPlan);
            const size_t len[3] = {(size_t)g.n[0], (size_t)g.n[1], (size_t)g.n[2]};
            QG_KFFT(rocfft_plan_create(&slot->fwd, rocfft_placement_notinplace, rocfft_transform_type_complex_forward,
                                       rocfft_precision_double, 3, len, nc, nullptr));
            QG_KFFT(rocfft_plan_create(&slot->inv, rocfft_placement_notinplace, rocfft_transform_type_complex_inverse,
                                       rocfft_precision_double, 3, len, nc, nullptr));
            QG_KFFT(rocfft_execution_info_create(&slot->info));
            size_t w1 = 0, w2 = 0;
            QG_KFFT(rocfft_plan_get_work_buffer_size(slot->fwd, &w1));
            QG_KFFT(rocfft_plan_get_work_buffer_size(slot->inv, &w2));
            slot->work.alloc(std::max<size_t>(1, std::max(w1, w2)));
        }
        QG_KFFT(rocfft_execution_info_set_stream(slot->info, s));
        QG_KFFT(rocfft_execution_info_set_work_buffer(slot->info, slot->work.p, slot->work.n));
        void* in[1] = {R + (int64_t)b0 * ng};
        void* out[1] = {kp->spec.p};
        QG_KFFT(rocfft_execute(slot->fwd, in, out, slot->info));
        hipLaunchKernelGGL(kinetic_scale_c2c_kernel, dim3((unsigned)std::min<int64_t>((ng * nc + 255) / 256, 8192)),
                           dim3(256), 0, s, kp->spec.p, g.n[0], g.n[1], g.n[2], nc, kp->tx.p, kp->ty.p, kp->tz.p, shift,
                           1.0 / (double)ng);
        QG_HIP(hipGetLastError());
        void* in2[1] = {kp->spec.p};
        void* out2[1] = {R + (int64_t)b0 * ng};
        QG_KFFT(rocfft_execute(slot->inv, in2, out2, slot->info));
    }
}

// planar FP64 bands [re(ng) | im(ng)] (the reference's SoAComplex blocks) <-> interleaved complex
template <class C>
__global__ void planar_to_z_kernel(int64_t ng, int nb, const double* __restrict__ planar, C* __restrict__ z) {
    const int64_t n = ng * nb;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / ng, i = t - b * ng;
        const double* p = planar + 2 * b * ng;
        z[t].x = (decltype(z[t].x))p[i];
        z[t].y = (decltype(z[t].y))p[ng + i];
    }
}

template <class C>
__global__ void z_to_planar_kernel(int64_t ng, int nb, const C* __restrict__ z, double* __restrict__ planar) {
    const int64_t n = ng * nb;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / ng, i = t - b * ng;
        double* p = planar + 2 * b * ng;
        p[i] = (double)z[t].x;
        p[ng + i] = (double)z[t].y;
    }
}

// ------------------------------------------------------------------ LOBPCG (complex FP64)
using zc = rocblas_double_complex;

__global__ void resid_kernel(const double2* AX, const double2* X, const double* theta, double2* R, int64_t K, int N) {
    const int64_t n = K * N;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(t / K);
        const double th = theta[b];
        R[t] = double2{AX[t].x - th * X[t].x, AX[t].y - th * X[t].y};
    }
}

// per column: sum |v|^2 (one workgroup per column)
__global__ void colnorm2_kernel(const double2* V, int64_t K, double* out) {
    __shared__ double red[kThreads / 64];
    const double2* v = V + (int64_t)blockIdx.x * K;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < K; i += kThreads) s += v[i].x * v[i].x + v[i].y * v[i].y;
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) t += red[w];
        out[blockIdx.x] = t;
    }
}

// SVQB's scaling of the Gram matrix G (n x n, in place) to unit diagonal: d_i = G_ii^-1/2 (0 for a
// zero column), G <- D G D.  Residual columns span many orders of magnitude as bands converge;
// the eigensolve and the relative drop threshold see the scaled matrix.  One workgroup.
__global__ void svqb_scale_kernel(double2* G, int n, double* d) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double gii = G[(int64_t)i * n + i].x;
        d[i] = gii > 0.0 ? 1.0 / sqrt(gii) : 0.0;
    }
    __syncthreads();
    for (int64_t t = threadIdx.x; t < (int64_t)n * n; t += blockDim.x) {
        const int i = (int)(t % n), j = (int)(t / n);
        const double sc = d[i] * d[j];
        G[t].x *= sc;
        G[t].y *= sc;
    }
}

// T[:, dst] = D U[:, j] / sqrt(lambda_j) for the kept eigenvalues (lambda_j > tau lambda_max),
// packed to the front largest first (dst = kept columns above j); one workgroup per source
// column; workgroup 0 writes the kept count
__global__ void svqb_pack_kernel(const double2* U, const double* lam, const double* d, int n, double tau, double2* T,
                                 int* keep) {
    const int j = blockIdx.x;
    const double lmax = lam[n - 1];
    auto kept = [&](int c) { return lam[c] > tau * lmax && lam[c] > 0.0; };
    int dst = 0;
    for (int c = n - 1; c > j; --c) dst += kept(c) ? 1 : 0;
    if (j == 0 && threadIdx.x == 0) {
        int m = 0;
        for (int c = 0; c < n; ++c) m += kept(c) ? 1 : 0;
        keep[0] = m;
    }
    if (!kept(j)) return;
    const double sc = 1.0 / sqrt(lam[j]);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double2 u = U[(int64_t)j * n + i];
        T[(int64_t)dst * n + i] = double2{u.x * sc * d[i], u.y * sc * d[i]};
    }
}

constexpr int kGramChunks = 64;

// G[t] = sum over chunks of part[c][t]
__global__ void sum_partials_kernel(const double2* __restrict__ part, int nc, int64_t mn, double2* __restrict__ G) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < mn; t += (int64_t)gridDim.x * blockDim.x) {
        double2 acc{0.0, 0.0};
        for (int c = 0; c < nc; ++c) {
            const double2 v = part[(int64_t)c * mn + t];
            acc.x += v.x;
            acc.y += v.y;
        }
        G[t] = acc;
    }
}

struct KLobpcg {
    qg_kpoint* kp;
    rocblas_handle hb;
    hipStream_t s;
    int64_t K;
    double dv;
    DevBuf<double2> U, Gm, Tm;   // small matrices
    DevBuf<double> lam, E, dsc;
    DevBuf<int> info, keep;   // rocSOLVER status; kept-column count of SVQB

    void gemm(rocblas_operation ta, rocblas_operation tb, int m, int n, int k, double alpha, const double2* A, int64_t lda,
              const double2* B, int64_t ldb, double beta, double2* Cm, int64_t ldc) {
        const zc al{alpha, 0.0}, be{beta, 0.0};
        QG_KBLAS(rocblas_zgemm(hb, ta, tb, m, n, k, &al, (const zc*)A, (int)lda, (const zc*)B, (int)ldb, &be, (zc*)Cm,
                               (int)ldc));
    }
    // G (m x n) = dV A^H B, A: K x m, B: K x n.  The products are skinny (m, n <= 3N, K = Ng), so
    // K is split into chunks, each chunk's product one batch entry of a strided-batched zgemm,
    // and the partial products summed: a single zgemm would keep only a few workgroups busy.
    DevBuf<double2> part;
    void gram(const double2* A, int m, const double2* B, int n, double2* G) {
        const int64_t kc = std::max<int64_t>(4096, (K + kGramChunks - 1) / kGramChunks);
        const int nc = (int)((K + kc - 1) / kc);
        const int64_t last = K - (int64_t)(nc - 1) * kc;
        const int64_t mn = (int64_t)m * n;
        part.alloc((size_t)nc * mn);
        const zc al{dv, 0.0}, be{0.0, 0.0};
        if (nc > 1)
            QG_KBLAS(rocblas_zgemm_strided_batched(hb, rocblas_operation_conjugate_transpose, rocblas_operation_none, m, n,
                                                   (int)kc, &al, (const zc*)A, (int)K, kc, (const zc*)B, (int)K, kc,
                                                   &be, (zc*)part.p, m, mn, nc - 1));
        // the last chunk (ragged) as one more product into its partial slot
        QG_KBLAS(rocblas_zgemm(hb, rocblas_operation_conjugate_transpose, rocblas_operation_none, m, n, (int)last, &al,
                               (const zc*)(A + (nc - 1) * kc), (int)K, (const zc*)(B + (nc - 1) * kc), (int)K, &be,
                               (zc*)(part.p + (int64_t)(nc - 1) * mn), m));
        hipLaunchKernelGGL(sum_partials_kernel, dim3((unsigned)std::min<int64_t>((mn + 255) / 256, 4096)), dim3(256), 0,
                           s, part.p, nc, mn, G);
        QG_HIP(hipGetLastError());
    }
    int heevd(double2* A, int n, double* w) {
        E.alloc(n);
        QG_HIP(hipMemsetAsync(info.p, 0, sizeof(int), s));
        QG_KBLAS(rocsolver_zheevd(hb, rocblas_evect_original, rocblas_fill_upper, n, (zc*)A, n, w, E.p, info.p));
        int inf = 0;
        QG_HIP(hipMemcpyAsync(&inf, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
        QG_HIP(hipStreamSynchronize(s));
        return inf;
    }
    void apply(const double2* in, double2* out, int n) {
        if (n > 0) kpoint_apply<double>(kp, n, in, K, out, K, s);
    }
    // V (K x n) <- V - Q (dV Q^H V), twice; AV alongside when given
    void project_out(const double2* Q, const double2* AQ, int q, double2* V, double2* AV, int n) {
        if (q == 0 || n == 0) return;
        for (int pass = 0; pass < 2; ++pass) {
            Gm.alloc((size_t)q * n);
            gram(Q, q, V, n, Gm.p);
            gemm(rocblas_operation_none, rocblas_operation_none, (int)K, n, q, -1.0, Q, K, Gm.p, q, 1.0, V, K);
            if (AV) gemm(rocblas_operation_none, rocblas_operation_none, (int)K, n, q, -1.0, AQ, K, Gm.p, q, 1.0, AV, K);
        }
    }
    // orthonormalise the n columns of V (and AV alongside) by SVQB on the unit-diagonal scaled
    // Gram matrix, dropping directions below tau of the largest; the kept columns are packed to
    // the front of V (and of AV).  Work: W (K x n).  Returns the number kept.
    int svqb(double2* V, double2* AV, int n, double2* W) {
        if (n == 0) return 0;
        Gm.alloc((size_t)n * n);
        U.alloc((size_t)n * n);
        Tm.alloc((size_t)n * n);
        lam.alloc(n);
        dsc.alloc(n);
        gram(V, n, V, n, U.p);
        hipLaunchKernelGGL(svqb_scale_kernel, dim3(1), dim3(256), 0, s, U.p, n, dsc.p);
        const int inf = heevd(U.p, n, lam.p);
        QG_REQUIRE(inf == 0, "k-point LOBPCG: Gram eigensolve failed (zheevd info " + std::to_string(inf) +
                                 ", n " + std::to_string(n) + ")");
        keep.alloc(1);
        hipLaunchKernelGGL(svqb_pack_kernel, dim3(n), dim3(256), 0, s, U.p, lam.p, dsc.p, n, 1e-13, Tm.p, keep.p);
        int m = 0;
        QG_HIP(hipMemcpyAsync(&m, keep.p, sizeof(int), hipMemcpyDeviceToHost, s));
        QG_HIP(hipStreamSynchronize(s));
        if (m == 0) return 0;
        gemm(rocblas_operation_none, rocblas_operation_none, (int)K, m, n, 1.0, V, K, Tm.p, n, 0.0, W, K);
        QG_HIP(hipMemcpyAsync(V, W, (size_t)K * m * sizeof(double2), hipMemcpyDeviceToDevice, s));
        if (AV) {
            gemm(rocblas_operation_none, rocblas_operation_none, (int)K, m, n, 1.0, AV, K, Tm.p, n, 0.0, W, K);
            QG_HIP(hipMemcpyAsync(AV, W, (size_t)K * m * sizeof(double2), hipMemcpyDeviceToDevice, s));
        }
        return m;
    }
};

int kpoint_lobpcg(qg_kpoint* kp, int N, double2* X, int64_t ld, double* eig_host, int iter_max, double tol, int nconv,
                  double precond_shift, int* iters_out, double* max_res_out, hipStream_t s) {
    qg_ham* h = kp->h;
    const int64_t K = h->grid.ng();
    QG_REQUIRE(ld == K, "k-point LOBPCG: X must be contiguous bands (ld = Ng)");
    QG_REQUIRE(N >= 1 && 3 * (int64_t)N <= K, "k-point LOBPCG: 1 <= nbands <= Ng / 3");
    if (nconv <= 0 || nconv > N) nconv = N;
    if (!kp->blas) QG_KBLAS(rocblas_create_handle(&kp->blas));
    QG_KBLAS(rocblas_set_stream(kp->blas, s));
    KLobpcg L{kp, kp->blas, s, K, h->grid.dvol()};
    L.info.alloc(1);
    const size_t blk = (size_t)K * N;
    // S = [X W P] and AS = [AX AW AP], contiguous so that one product forms the Rayleigh-Ritz
    // matrix; T1, T2 work blocks
    DevBuf<double2> S, AS, T1, T2, T3;
    S.alloc(3 * blk);
    AS.alloc(3 * blk);
    T1.alloc(blk);
    T2.alloc(blk);
    T3.alloc(blk);   // AP while P is orthonormalised
    double2* Sx = S.p;
    double2* ASx = AS.p;
    QG_HIP(hipMemcpyAsync(Sx, X, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
    // X orthonormal, then the Rayleigh-Ritz of X alone
    int mx = L.svqb(Sx, nullptr, N, T1.p);
    QG_REQUIRE(mx == N, "k-point LOBPCG: the initial block is rank deficient");
    L.apply(Sx, ASx, N);
    DevBuf<double2> A, Cm;
    DevBuf<double> theta, nrm;
    A.alloc((size_t)9 * N * N);
    theta.alloc(3 * N);
    nrm.alloc(N);
    L.gram(Sx, N, ASx, N, A.p);
    const int inf0 = L.heevd(A.p, N, theta.p);
    QG_REQUIRE(inf0 == 0, "k-point LOBPCG: Rayleigh-Ritz eigensolve failed (zheevd info " + std::to_string(inf0) + ")");
    L.gemm(rocblas_operation_none, rocblas_operation_none, (int)K, N, N, 1.0, Sx, K, A.p, N, 0.0, T1.p, K);
    L.gemm(rocblas_operation_none, rocblas_operation_none, (int)K, N, N, 1.0, ASx, K, A.p, N, 0.0, T2.p, K);
    QG_HIP(hipMemcpyAsync(Sx, T1.p, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
    QG_HIP(hipMemcpyAsync(ASx, T2.p, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
    int np = 0;   // columns of P in use
    std::vector<double> th(N), r2(N);
    double maxres = 0.0;
    int it = 0;
    const unsigned rg = (unsigned)std::min<int64_t>((K * N + 255) / 256, 65536);
    for (;; ++it) {
        double2* Sw = Sx + (int64_t)K * N;    // W slot
        double2* ASw = ASx + (int64_t)K * N;
        // residuals R = AX - X theta into T1
        hipLaunchKernelGGL(resid_kernel, dim3(rg), dim3(256), 0, s, ASx, Sx, theta.p, T1.p, K, N);
        hipLaunchKernelGGL(colnorm2_kernel, dim3(N), dim3(kThreads), 0, s, T1.p, K, nrm.p);
        QG_HIP(hipMemcpyAsync(r2.data(), nrm.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
        QG_HIP(hipMemcpyAsync(th.data(), theta.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
        QG_HIP(hipStreamSynchronize(s));
        maxres = 0.0;
        for (int b = 0; b < nconv; ++b) maxres = std::max(maxres, std::sqrt(r2[b] * L.dv));
        if (maxres <= tol || it >= iter_max) break;
        // W: the residuals with X (and P) projected out, orthonormalised
        // P occupies the slot after the kept W columns; move it out of the way first
        double2* Pcur = T2.p;
        double2* APcur = T3.p;
        if (np > 0) {
            QG_HIP(hipMemcpyAsync(Pcur, Sx + (int64_t)K * 2 * N, (size_t)K * np * sizeof(double2),
                                  hipMemcpyDeviceToDevice, s));
            QG_HIP(hipMemcpyAsync(APcur, ASx + (int64_t)K * 2 * N, (size_t)K * np * sizeof(double2),
                                  hipMemcpyDeviceToDevice, s));
        }
        QG_HIP(hipMemcpyAsync(Sw, T1.p, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
        if (precond_shift > 0.0) kinetic_precond(kp, N, Sw, precond_shift, s);   // W = K R
        L.project_out(Sx, nullptr, N, Sw, nullptr, N);
        const int mw = L.svqb(Sw, nullptr, N, T1.p);
        L.apply(Sw, ASw, mw);
        // P: projected out of X and W, orthonormalised, placed after W
        int mp = 0;
        if (np > 0) {
            L.project_out(Sx, ASx, N, Pcur, APcur, np);
            L.project_out(Sw, ASw, mw, Pcur, APcur, np);
            mp = L.svqb(Pcur, APcur, np, T1.p);
            double2* Sp = Sw + (int64_t)K * mw;
            double2* ASp = ASw + (int64_t)K * mw;
            if (mp > 0) {
                QG_HIP(hipMemcpyAsync(Sp, Pcur, (size_t)K * mp * sizeof(double2), hipMemcpyDeviceToDevice, s));
                QG_HIP(hipMemcpyAsync(ASp, APcur, (size_t)K * mp * sizeof(double2), hipMemcpyDeviceToDevice, s));
            }
        }
        const int m = N + mw + mp;
        // Rayleigh-Ritz on the orthonormal basis S (m columns)
        L.gram(Sx, m, ASx, m, A.p);
        const int infm = L.heevd(A.p, m, theta.p);
        QG_REQUIRE(infm == 0, "k-point LOBPCG: Rayleigh-Ritz eigensolve failed (zheevd info " + std::to_string(infm) +
                                  ", n " + std::to_string(m) + ")");
        // new P = S[:, N:m] C[N:m, :N], AP alike (into T1 / T2), then X = S C[:, :N]
        const int mr = m - N;
        if (mr > 0) {
            L.gemm(rocblas_operation_none, rocblas_operation_none, (int)K, N, mr, 1.0, Sw, K, A.p + N, m, 0.0, T1.p, K);
            L.gemm(rocblas_operation_none, rocblas_operation_none, (int)K, N, mr, 1.0, ASw, K, A.p + N, m, 0.0, T2.p, K);
        }
        // X <- X C_x + P_new, computed in the W slot (free now), then moved to the X slot
        if (mr > 0) {
            QG_HIP(hipMemcpyAsync(Sw, T1.p, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
            QG_HIP(hipMemcpyAsync(ASw, T2.p, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
            L.gemm(rocblas_operation_none, rocblas_operation_none, (int)K, N, N, 1.0, Sx, K, A.p, m, 1.0, Sw, K);
            L.gemm(rocblas_operation_none, rocblas_operation_none, (int)K, N, N, 1.0, ASx, K, A.p, m, 1.0, ASw, K);
        } else {
            L.gemm(rocblas_operation_none, rocblas_operation_none, (int)K, N, N, 1.0, Sx, K, A.p, m, 0.0, Sw, K);
            L.gemm(rocblas_operation_none, rocblas_operation_none, (int)K, N, N, 1.0, ASx, K, A.p, m, 0.0, ASw, K);
        }
        QG_HIP(hipMemcpyAsync(Sx, Sw, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
        QG_HIP(hipMemcpyAsync(ASx, ASw, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
        // P slot: the new P (N columns) from T1 / T2
        np = mr > 0 ? N : 0;
        if (np > 0) {
            QG_HIP(hipMemcpyAsync(Sx + (int64_t)K * 2 * N, T1.p, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
            QG_HIP(hipMemcpyAsync(ASx + (int64_t)K * 2 * N, T2.p, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
        }
        QG_HIP(hipStreamSynchronize(s));
    }
    QG_HIP(hipMemcpyAsync(X, Sx, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
    QG_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < N; ++b) eig_host[b] = th[b];
    if (iters_out) *iters_out = it;
    if (max_res_out) *max_res_out = maxres;
    return QG_OK;
}

}  // namespace
}  // namespace qg

using namespace qg;

extern "C" {

int qg_kpoint_create(qg_ham* h, double kx, double ky, double kz, qg_kpoint** out) {
    return guarded([&] {
        QG_REQUIRE(h && out, "null pointer");
        qg_kpoint* kp = new qg_kpoint;
        kp->h = h;
        kp->k[0] = kx;
        kp->k[1] = ky;
        kp->k[2] = kz;
        try {
            build_phases(kp);
        } catch (...) {
            delete kp;
            throw;
        }
        *out = kp;
    });
}

int qg_kpoint_destroy(qg_kpoint* kp) {
    return guarded([&] { delete kp; });
}

int qg_kpoint_apply(qg_kpoint* kp, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi, int64_t ld_out,
                    void* stream) {
    return guarded([&] {
        check_kblock(kp, dtype, nbands, psi, ld, hpsi, ld_out);
        if (nbands == 0) return;
        hipStream_t s = as_stream(stream);
        if (dtype == QG_Z128)
            kpoint_apply<double>(kp, nbands, psi, ld, hpsi, ld_out, s);
        else
            kpoint_apply<float>(kp, nbands, psi, ld, hpsi, ld_out, s);
    });
}

int qg_complex_from_planar(int64_t ng, int nbands, const double* planar, int dtype, void* z, void* stream) {
    return guarded([&] {
        QG_REQUIRE(ng >= 0 && nbands >= 0, "negative size");
        QG_REQUIRE(dtype == QG_Z128 || dtype == QG_C64, "complex dtype required (QG_Z128 or QG_C64)");
        if (ng == 0 || nbands == 0) return;
        QG_REQUIRE(planar && z, "null pointer");
        const unsigned grid = (unsigned)std::min<int64_t>((ng * nbands + 255) / 256, 65536);
        if (dtype == QG_Z128)
            hipLaunchKernelGGL(planar_to_z_kernel<double2>, dim3(grid), dim3(256), 0, as_stream(stream), ng, nbands,
                               planar, (double2*)z);
        else
            hipLaunchKernelGGL(planar_to_z_kernel<float2>, dim3(grid), dim3(256), 0, as_stream(stream), ng, nbands,
                               planar, (float2*)z);
        QG_HIP(hipGetLastError());
    });
}

int qg_complex_to_planar(int64_t ng, int nbands, const void* z, int dtype, double* planar, void* stream) {
    return guarded([&] {
        QG_REQUIRE(ng >= 0 && nbands >= 0, "negative size");
        QG_REQUIRE(dtype == QG_Z128 || dtype == QG_C64, "complex dtype required (QG_Z128 or QG_C64)");
        if (ng == 0 || nbands == 0) return;
        QG_REQUIRE(planar && z, "null pointer");
        const unsigned grid = (unsigned)std::min<int64_t>((ng * nbands + 255) / 256, 65536);
        if (dtype == QG_Z128)
            hipLaunchKernelGGL(z_to_planar_kernel<double2>, dim3(grid), dim3(256), 0, as_stream(stream), ng, nbands,
                               (const double2*)z, planar);
        else
            hipLaunchKernelGGL(z_to_planar_kernel<float2>, dim3(grid), dim3(256), 0, as_stream(stream), ng, nbands,
                               (const float2*)z, planar);
        QG_HIP(hipGetLastError());
    });
}

int qg_kpoint_lobpcg(qg_kpoint* kp, int nbands, void* X, int64_t ld, double* eig_host, int iter_max, double tol,
                     int nconv, double precond_shift, int* iters_out, double* max_res_out, void* stream) {
    return guarded([&] {
        QG_REQUIRE(kp && kp->h && X && eig_host, "null pointer");
        QG_REQUIRE(iter_max >= 0, "negative iteration limit");
        kpoint_lobpcg(kp, nbands, (double2*)X, ld, eig_host, iter_max, tol, nconv, precond_shift, iters_out,
                      max_res_out, as_stream(stream));
    });
}

}  // extern "C"
