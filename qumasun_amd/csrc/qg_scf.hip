// qg_scf.hip — the self-consistent-field step around the eigensolver, device-resident on one
// GPU: occupancies, density, linear density mixing, the Hartree potential by an FFT Poisson
// solve (rocFFT), LDA exchange-correlation (Perdew-Zunger 1981) and the total local potential
// installed in the Hamiltonian.
//
// Reference semantics (QUMASUN source/):
//   occupancy  coef / (exp(min(beta (e - mu), 80)) + 1), coef 2 (no spin), mu by bisection on
//              [e_0, max] until the bracket is below 1e-9      SetOccupancy.h:7-18,
//                                                              qumasun_mpi.occupancy.h:61-108
//   density    rho = sum over bands with occ > 0 of occ psi^2  qumasun_mpi.density.h:10-32
//   mixing     rho = prev + ratio (rho - prev); int |rho - prev| dV   qumasun_mpi.density.h:37-67
//   Poisson    V_H(k) = 4 pi rho(k) / |k|^2, k_d = 2 pi m_d / L_d with m folded to
//              (m < n/2 ? m : m - n), V_H(0) = 0; inverse transform scaled by 1/N
//                                                              poisson_fft.h:131-240, wrap_fft.h:14-38
//   LDA        Calc_XC_LDA(rho + pcc): zero below rho = 1e-8; rs = (3/4pi)^(1/3) rho^(-1/3);
//              e_x = -(3/4pi)(9pi/4)^(1/3) / rs, V_x = 4/3 e_x; PZ81 correlation with the
//              rs >= 1 and rs < 1 branches                     Vxc.h:17-75
//   V_tot = V_ext + V_H + V_x + V_c                            qumasun_mpi.potential.h:140-145
#include <rocfft/rocfft.h>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>
#include "qg_internal.hpp"

namespace qg {

#define QG_FFT_CALL(call)                                                             \
    do {                                                                              \
        rocfft_status st_ = (call);                                                   \
        if (st_ != rocfft_status_success)                                             \
            throw ::qg::Error{QG_ERR_LIB, std::string(#call) + " failed: " + std::to_string((int)st_)}; \
    } while (0)

namespace {

constexpr int kThreads = 256;

int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 8192)); }

// rho[i] = sum_b occ_b psi_b[i]^2 over the bands with occ_b > 0 (in band order, FP64)
template <class T>
__global__ void density_kernel(const T* __restrict__ psi, int64_t ld, int nbands, const double* __restrict__ occ,
                               int64_t ng, double* __restrict__ rho) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ng; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nbands; ++b) {
            const double f = occ[b];
            if (f > 0.0) {
                const double p = (double)psi[(int64_t)b * ld + i];
                s += (p * p) * f;
            }
        }
        rho[i] = s;
    }
}

// rho <- prev + ratio (rho - prev); per-block partial sums of |rho - prev| and of the mixed rho
__global__ void mix_kernel(double* __restrict__ rho, const double* __restrict__ prev, int64_t ng, double ratio,
                           int mix, double* __restrict__ part) {
    __shared__ double sd[kThreads], sr[kThreads];
    double d = 0.0, r = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ng; i += (int64_t)gridDim.x * blockDim.x) {
        double v = rho[i];
        if (mix) {
            const double diff = v - prev[i];
            v = diff * ratio + prev[i];
            rho[i] = v;
            d += fabs(diff);
        }
        r += v;
    }
    sd[threadIdx.x] = d;
    sr[threadIdx.x] = r;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sd[threadIdx.x] += sd[threadIdx.x + w];
            sr[threadIdx.x] += sr[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sd[0];
        part[2 * blockIdx.x + 1] = sr[0];
    }
}

// the Poisson kernel on the half spectrum of a real-to-complex transform: 4 pi / |k|^2 / N
__global__ void poisson_scale_kernel(double2* __restrict__ spec, int nxh, int nx, int ny, int nz, double cx, double cy,
                                     double cz, double inv_n) {
    const int64_t total = (int64_t)nxh * ny * nz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i % nxh);
        const int ky = (int)((i / nxh) % ny);
        const int kz = (int)(i / ((int64_t)nxh * ny));
        const double fx = cx * (double)(kx < nx / 2 ? kx : kx - nx);
        const double fy = cy * (double)(ky < ny / 2 ? ky : ky - ny);
        const double fz = cz * (double)(kz < nz / 2 ? kz : kz - nz);
        const double k2 = fx * fx + fy * fy + fz * fz;
        const double f = (kx == 0 && ky == 0 && kz == 0) ? 0.0 : 4.0 * M_PI / k2 * inv_n;
        spec[i].x *= f;
        spec[i].y *= f;
    }
}

struct Xc {
    double ex, ec, vx, vc;
};

__device__ __forceinline__ Xc lda_pz81(double rho) {
    Xc o{0.0, 0.0, 0.0, 0.0};
    if (rho < 1.0e-8) return o;
    const double c2 = 0.62035049089940009;    // (3/4pi)^(1/3)
    const double c3 = -0.45816529328314287;   // -(3/4pi)(9pi/4)^(1/3)
    const double rs = c2 / cbrt(rho);
    const double ex = c3 / rs;
    double ec, ndec;
    if (rs >= 1.0) {
        const double g = -0.1423, b1 = 1.0529, b2 = 0.3334;
        const double sq = sqrt(rs);
        ec = g / (1.0 + b1 * sq + b2 * rs);
        ndec = ec * ec * (b1 * 0.5 * sq + b2 * rs) / (3.0 * g);
    } else {
        const double A = 0.0311, B = -0.0480, C = 0.0020, D = -0.0116;
        const double lr = log(rs);
        ec = B + A * lr + D * rs + C * rs * lr;
        ndec = -(A + (D + C) * rs + C * rs * lr) / 3.0;
    }
    o.ex = ex;
    o.ec = ec;
    o.vx = ex + ex / 3.0;
    o.vc = ec + ndec;
    return o;
}

__global__ void lda_kernel(const double* __restrict__ rho, const double* __restrict__ pcc, int64_t ng,
                           double* __restrict__ vx, double* __restrict__ vc, double* __restrict__ ex,
                           double* __restrict__ ec) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ng; i += (int64_t)gridDim.x * blockDim.x) {
        const Xc r = lda_pz81(rho[i] + (pcc ? pcc[i] : 0.0));
        vx[i] = r.vx;
        vc[i] = r.vc;
        if (ex) ex[i] = r.ex;
        if (ec) ec[i] = r.ec;
    }
}

// V = a + b + c + d (any of b..d may be absent); the FP32 copy alongside
__global__ void vtot_kernel(const double* __restrict__ vext, const double* __restrict__ vh, const double* __restrict__ vx,
                            const double* __restrict__ vc, int64_t ng, double* __restrict__ v64, float* __restrict__ v32,
                            double* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ng; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = (vext ? vext[i] : 0.0) + (vh ? vh[i] : 0.0) + (vx ? vx[i] : 0.0) + (vc ? vc[i] : 0.0);
        v64[i] = v;
        v32[i] = (float)v;
        if (out) out[i] = v;
    }
}

__global__ void narrow_copy_kernel(const double* __restrict__ v, int64_t ng, double* __restrict__ v64,
                                   float* __restrict__ v32) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ng; i += (int64_t)gridDim.x * blockDim.x) {
        v64[i] = v[i];
        v32[i] = (float)v[i];
    }
}

// FP64 real-to-complex / complex-to-real plans of one grid shape, kept for the process
struct PoissonPlan {
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    DevBuf<unsigned char> work;
    DevBuf<double2> spec;
    ~PoissonPlan() {
        if (fwd) rocfft_plan_destroy(fwd);
        if (inv) rocfft_plan_destroy(inv);
        if (info) rocfft_execution_info_destroy(info);
    }
};

std::mutex g_plan_mu;

PoissonPlan& poisson_plan(const qg_grid& g) {
    static std::map<std::tuple<int, int, int, int>, std::unique_ptr<PoissonPlan>> plans;
    static std::once_flag once;
    std::call_once(once, [] { rocfft_setup(); });
    int dev = 0;
    QG_HIP(hipGetDevice(&dev));
    auto& p = plans[std::make_tuple(dev, g.n[0], g.n[1], g.n[2])];
    if (!p) {
        auto q = std::make_unique<PoissonPlan>();
        const size_t len[3] = {(size_t)g.n[0], (size_t)g.n[1], (size_t)g.n[2]};
        QG_FFT_CALL(rocfft_plan_create(&q->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                       rocfft_precision_double, 3, len, 1, nullptr));
        QG_FFT_CALL(rocfft_plan_create(&q->inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                                       rocfft_precision_double, 3, len, 1, nullptr));
        QG_FFT_CALL(rocfft_execution_info_create(&q->info));
        size_t w1 = 0, w2 = 0;
        QG_FFT_CALL(rocfft_plan_get_work_buffer_size(q->fwd, &w1));
        QG_FFT_CALL(rocfft_plan_get_work_buffer_size(q->inv, &w2));
        q->work.alloc(std::max<size_t>(1, std::max(w1, w2)));
        q->spec.alloc((size_t)(g.n[0] / 2 + 1) * g.n[1] * g.n[2]);
        p = std::move(q);
    }
    return *p;
}

}  // namespace

void hartree_potential(const qg_grid& g, const double* rho, double* vh, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    PoissonPlan& pl = poisson_plan(g);
    QG_FFT_CALL(rocfft_execution_info_set_stream(pl.info, s));
    QG_FFT_CALL(rocfft_execution_info_set_work_buffer(pl.info, pl.work.p, pl.work.n));
    void* in[1] = {const_cast<double*>(rho)};
    void* out[1] = {pl.spec.p};
    QG_FFT_CALL(rocfft_execute(pl.fwd, in, out, pl.info));
    const int nxh = g.n[0] / 2 + 1;
    const int64_t nspec = (int64_t)nxh * g.n[1] * g.n[2];
    hipLaunchKernelGGL(poisson_scale_kernel, dim3(blocks_for(nspec)), dim3(kThreads), 0, s, pl.spec.p, nxh, g.n[0],
                       g.n[1], g.n[2], 2.0 * M_PI / (g.h[0] * g.n[0]), 2.0 * M_PI / (g.h[1] * g.n[1]),
                       2.0 * M_PI / (g.h[2] * g.n[2]), 1.0 / (double)g.ng());
    QG_HIP(hipGetLastError());
    void* in2[1] = {pl.spec.p};
    void* out2[1] = {vh};
    QG_FFT_CALL(rocfft_execute(pl.inv, in2, out2, pl.info));
    // the plans' buffers are shared by the grid's callers: finish before releasing them
    QG_HIP(hipStreamSynchronize(s));
}

}  // namespace qg

using namespace qg;

extern "C" {

int qg_occupancy_fermi(int nbands, const double* eig, double num_electrons, double kbT, double* occ, double* mu_out) {
    return guarded([&] {
        QG_REQUIRE(nbands >= 1 && eig && occ, "bad arguments");
        QG_REQUIRE(kbT > 0.0, "kbT must be positive");
        const double beta = 1.0 / kbT;
        auto fill = [&](double mu) {
            double sum = 0.0;
            for (int s = 0; s < nbands; ++s) {
                const double p = std::min(beta * (eig[s] - mu), 80.0);
                occ[s] = 2.0 / (std::exp(p) + 1.0);
                sum += occ[s];
            }
            return sum;
        };
        double mu_min = eig[0];
        double mu_max = eig[nbands - 1];
        mu_max = (mu_max > 0.0) ? mu_max * 16.0 : std::fabs(mu_max) + 16.0;
        // only the minimum number of states solved: a wider upper bracket
        if (((int64_t)std::llround(num_electrons) + 1) / 2 >= nbands) mu_max *= 10.0;
        double mu = (mu_min + mu_max) / 2.0;
        fill(mu);
        while (std::fabs(mu_max - mu_min) > 1.0e-9) {
            const double ne = fill(mu);
            if (ne < num_electrons) mu_min = mu;
            else if (ne > num_electrons) mu_max = mu;
            else break;
            mu = (mu_min + mu_max) / 2.0;
        }
        if (mu_out) *mu_out = mu;
    });
}

int qg_density(const qg_grid* g, int dtype, int nbands, const void* psi, int64_t ld, const double* occ_host,
               double* rho, void* stream) {
    return guarded([&] {
        QG_REQUIRE(g && rho && (nbands == 0 || (psi && occ_host)), "null argument");
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        QG_REQUIRE(nbands >= 0 && ld >= g->ng(), "bad band block");
        hipStream_t s = as_stream(stream);
        DevBuf<double> occ;
        occ.upload(occ_host, std::max(1, nbands));
        const int64_t ng = g->ng();
        if (dtype == QG_F64)
            hipLaunchKernelGGL(density_kernel<double>, dim3(blocks_for(ng)), dim3(kThreads), 0, s, (const double*)psi, ld,
                               nbands, occ.p, ng, rho);
        else
            hipLaunchKernelGGL(density_kernel<float>, dim3(blocks_for(ng)), dim3(kThreads), 0, s, (const float*)psi, ld,
                               nbands, occ.p, ng, rho);
        QG_HIP(hipGetLastError());
        QG_HIP(hipStreamSynchronize(s));
    });
}

int qg_density_mix(const qg_grid* g, double* rho, const double* rho_prev, double ratio, double* diff_abs,
                   double* total, void* stream) {
    return guarded([&] {
        QG_REQUIRE(g && rho, "null argument");
        hipStream_t s = as_stream(stream);
        const int64_t ng = g->ng();
        const int nb = blocks_for(ng);
        DevBuf<double> part;
        part.alloc(2 * (size_t)nb);
        hipLaunchKernelGGL(mix_kernel, dim3(nb), dim3(kThreads), 0, s, rho, rho_prev, ng, ratio, rho_prev ? 1 : 0,
                           part.p);
        QG_HIP(hipGetLastError());
        std::vector<double> h(2 * (size_t)nb);
        QG_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * 8, hipMemcpyDeviceToHost, s));
        QG_HIP(hipStreamSynchronize(s));
        double d = 0.0, r = 0.0;
        for (int b = 0; b < nb; ++b) {
            d += h[2 * b];
            r += h[2 * b + 1];
        }
        if (diff_abs) *diff_abs = d * g->dvol();
        if (total) *total = r * g->dvol();
    });
}

int qg_hartree_potential(const qg_grid* g, const double* rho, double* vhart, void* stream) {
    return guarded([&] {
        QG_REQUIRE(g && rho && vhart, "null argument");
        hartree_potential(*g, rho, vhart, as_stream(stream));
    });
}

int qg_xc_lda(const qg_grid* g, const double* rho, const double* pcc, double* vx, double* vc, double* ex, double* ec,
              void* stream) {
    return guarded([&] {
        QG_REQUIRE(g && rho && vx && vc, "null argument");
        hipStream_t s = as_stream(stream);
        const int64_t ng = g->ng();
        hipLaunchKernelGGL(lda_kernel, dim3(blocks_for(ng)), dim3(kThreads), 0, s, rho, pcc, ng, vx, vc, ex, ec);
        QG_HIP(hipGetLastError());
        QG_HIP(hipStreamSynchronize(s));
    });
}

int qg_hamiltonian_set_potential_device(qg_ham* h, const double* V, void* stream) {
    return guarded([&] {
        QG_REQUIRE(h && V, "null argument");
        hipStream_t s = as_stream(stream);
        const int64_t ng = h->grid.ng();
        hipLaunchKernelGGL(narrow_copy_kernel, dim3(blocks_for(ng)), dim3(kThreads), 0, s, V, ng, h->V64.p, h->V32.p);
        QG_HIP(hipGetLastError());
        QG_HIP(hipStreamSynchronize(s));
    });
}

int qg_scf_potential(qg_ham* h, const double* rho, const double* pcc, const double* vext, double* vhart, double* vx,
                     double* vc, double* vtot, void* stream) {
    return guarded([&] {
        QG_REQUIRE(h && rho && vhart && vx && vc, "null argument");
        hipStream_t s = as_stream(stream);
        const qg_grid& g = h->grid;
        const int64_t ng = g.ng();
        hartree_potential(g, rho, vhart, s);
        hipLaunchKernelGGL(lda_kernel, dim3(blocks_for(ng)), dim3(kThreads), 0, s, rho, pcc, ng, vx, vc,
                           (double*)nullptr, (double*)nullptr);
        QG_HIP(hipGetLastError());
        hipLaunchKernelGGL(vtot_kernel, dim3(blocks_for(ng)), dim3(kThreads), 0, s, vext, (const double*)vhart,
                           (const double*)vx, (const double*)vc, ng, h->V64.p, h->V32.p, vtot);
        QG_HIP(hipGetLastError());
        QG_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
