// qg_precond.hip — preconditioners for the LOBPCG residual block.
//
// The reference uses the raw residual (lobpcg_d_multi_mpi.h:982-990, :1372-1378).  This stage
// adds the kinetic preconditioner K = (T + s)^-1 with T = -1/2 Lap8, the periodic 8th-order
// finite-difference kinetic operator.  On a periodic grid T is diagonal in the discrete
// Fourier basis with the symbol
//   T(k) = -1/2 sum_d h_d^-2 [ w0 + 2 sum_{m=1..4} w_m cos(m theta_d) ],  theta_d = 2 pi k_d / n_d,
// (w0, w1..w4) = (-205/72, 8/5, -1/5, 8/315, -1/560), so K is applied exactly with one real-to-
// complex FFT, a pointwise scale by 1 / (T(k) + s) / Ng, and one complex-to-real FFT (rocFFT),
// batched over bands in chunks.
#include <rocfft/rocfft.h>
#include <cmath>
#include <mutex>
#include <vector>
#include "qg_internal.hpp"

namespace qg {

#define QG_FFT(call)                                                                  \
    do {                                                                              \
        rocfft_status st_ = (call);                                                   \
        if (st_ != rocfft_status_success)                                             \
            throw ::qg::Error{QG_ERR_LIB, std::string(#call) + " failed: " + std::to_string((int)st_)}; \
    } while (0)

struct PrecondPlan {
    int n[3] = {0, 0, 0};
    int prec = -1;
    int batch = 0;
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    DevBuf<unsigned char> work, spec;
    DevBuf<double> tx, ty, tz;   // 1-D symbol parts
    ~PrecondPlan() {
        if (fwd) rocfft_plan_destroy(fwd);
        if (inv) rocfft_plan_destroy(inv);
        if (info) rocfft_execution_info_destroy(info);
    }
};

// 1-D part of the kinetic symbol: -1/2 h^-2 [w0 + 2 sum_m w_m cos(m theta)]
static std::vector<double> kinetic_symbol_1d(int n, double h, int count) {
    const double w[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
    std::vector<double> t(count);
    for (int k = 0; k < count; ++k) {
        const double th = 2.0 * M_PI * k / n;
        double s = w[0];
        for (int m = 1; m <= 4; ++m) s += 2.0 * w[m] * std::cos(m * th);
        t[k] = -0.5 * s / (h * h);
    }
    return t;
}

template <class C>
__global__ void kinetic_scale_kernel(C* spec, int nxh, int ny, int nz, int nb, const double* tx, const double* ty,
                                     const double* tz, double shift, double inv_ng) {
    const int64_t per = (int64_t)nxh * ny * nz;
    const int64_t total = per * nb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i % per;
        const int kx = (int)(j % nxh);
        const int ky = (int)((j / nxh) % ny);
        const int kz = (int)(j / ((int64_t)nxh * ny));
        const double f = inv_ng / (tx[kx] + ty[ky] + tz[kz] + shift);
        spec[i].x = (decltype(spec[i].x))(spec[i].x * f);
        spec[i].y = (decltype(spec[i].y))(spec[i].y * f);
    }
}

static void ensure_setup() {
    static std::once_flag once;
    std::call_once(once, [] { rocfft_setup(); });
}

void precond_apply(qg_ham* h, int kind, double shift, int dtype, int nbands, void* r, int64_t ld, hipStream_t s) {
    QG_REQUIRE(kind == QG_PRECOND_KINETIC, "unknown preconditioner");
    QG_REQUIRE(shift > 0.0, "the kinetic preconditioner needs a positive shift");
    QG_REQUIRE(ld == h->grid.ng(), "preconditioner expects ld == Ng");
    ensure_setup();
    const qg_grid& g = h->grid;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int nxh = nx / 2 + 1;
    const size_t es = dtype == QG_F32 ? 4 : 8;
    const int64_t ng = g.ng();
    const int64_t nspec = (int64_t)nxh * ny * nz;
    // bands per FFT batch: about 256 MB of spectrum
    const int batch = (int)std::max<int64_t>(1, std::min<int64_t>(nbands, (256ll << 20) / (nspec * 2 * es)));
    if (!h->precond) h->precond = new PrecondPlan;
    PrecondPlan& pl = *h->precond;
    if (pl.n[0] != nx || pl.n[1] != ny || pl.n[2] != nz || pl.prec != dtype || pl.batch != batch) {
        if (pl.fwd) { rocfft_plan_destroy(pl.fwd); pl.fwd = nullptr; }
        if (pl.inv) { rocfft_plan_destroy(pl.inv); pl.inv = nullptr; }
        if (!pl.info) QG_FFT(rocfft_execution_info_create(&pl.info));
        const size_t len[3] = {(size_t)nx, (size_t)ny, (size_t)nz};
        const rocfft_precision p = dtype == QG_F32 ? rocfft_precision_single : rocfft_precision_double;
        QG_FFT(rocfft_plan_create(&pl.fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward, p, 3, len,
                                  batch, nullptr));
        QG_FFT(rocfft_plan_create(&pl.inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse, p, 3, len,
                                  batch, nullptr));
        size_t w1 = 0, w2 = 0;
        QG_FFT(rocfft_plan_get_work_buffer_size(pl.fwd, &w1));
        QG_FFT(rocfft_plan_get_work_buffer_size(pl.inv, &w2));
        pl.work.alloc(std::max<size_t>(1, std::max(w1, w2)));
        pl.spec.alloc((size_t)nspec * batch * 2 * es);
        auto a = kinetic_symbol_1d(nx, g.h[0], nxh);
        auto b = kinetic_symbol_1d(ny, g.h[1], ny);
        auto c = kinetic_symbol_1d(nz, g.h[2], nz);
        pl.tx.upload(a.data(), a.size());
        pl.ty.upload(b.data(), b.size());
        pl.tz.upload(c.data(), c.size());
        pl.n[0] = nx; pl.n[1] = ny; pl.n[2] = nz;
        pl.prec = dtype;
        pl.batch = batch;
    }
    QG_FFT(rocfft_execution_info_set_stream(pl.info, s));
    QG_FFT(rocfft_execution_info_set_work_buffer(pl.info, pl.work.p, pl.work.n));
    for (int b0 = 0; b0 < nbands; b0 += batch) {
        // rocFFT plans have a fixed batch: the last chunk re-transforms a few bands of the
        // previous chunk (idempotent only for the forward part), so process it separately
        const int nb = std::min(batch, nbands - b0);
        unsigned char* base = (unsigned char*)r + (size_t)b0 * ng * es;
        if (nb != batch) {
            // shift the window back so a full batch is transformed; the overlap bands are
            // saved first and restored afterwards
            const int back = batch - nb;
            unsigned char* wbase = base - (size_t)back * ng * es;
            DevBuf<unsigned char> save;
            save.alloc((size_t)back * ng * es);
            QG_HIP(hipMemcpyAsync(save.p, wbase, (size_t)back * ng * es, hipMemcpyDeviceToDevice, s));
            void* in[1] = {wbase};
            void* out[1] = {pl.spec.p};
            QG_FFT(rocfft_execute(pl.fwd, in, out, pl.info));
            if (dtype == QG_F32)
                hipLaunchKernelGGL(kinetic_scale_kernel<float2>, dim3(4096), dim3(256), 0, s, (float2*)pl.spec.p, nxh, ny,
                                   nz, batch, pl.tx.p, pl.ty.p, pl.tz.p, shift, 1.0 / (double)ng);
            else
                hipLaunchKernelGGL(kinetic_scale_kernel<double2>, dim3(4096), dim3(256), 0, s, (double2*)pl.spec.p, nxh,
                                   ny, nz, batch, pl.tx.p, pl.ty.p, pl.tz.p, shift, 1.0 / (double)ng);
            void* in2[1] = {pl.spec.p};
            void* out2[1] = {wbase};
            QG_FFT(rocfft_execute(pl.inv, in2, out2, pl.info));
            QG_HIP(hipMemcpyAsync(wbase, save.p, (size_t)back * ng * es, hipMemcpyDeviceToDevice, s));
            QG_HIP(hipStreamSynchronize(s));
            continue;
        }
        void* in[1] = {base};
        void* out[1] = {pl.spec.p};
        QG_FFT(rocfft_execute(pl.fwd, in, out, pl.info));
        if (dtype == QG_F32)
            hipLaunchKernelGGL(kinetic_scale_kernel<float2>, dim3(4096), dim3(256), 0, s, (float2*)pl.spec.p, nxh, ny, nz,
                               batch, pl.tx.p, pl.ty.p, pl.tz.p, shift, 1.0 / (double)ng);
        else
            hipLaunchKernelGGL(kinetic_scale_kernel<double2>, dim3(4096), dim3(256), 0, s, (double2*)pl.spec.p, nxh, ny,
                               nz, batch, pl.tx.p, pl.ty.p, pl.tz.p, shift, 1.0 / (double)ng);
        void* in2[1] = {pl.spec.p};
        void* out2[1] = {base};
        QG_FFT(rocfft_execute(pl.inv, in2, out2, pl.info));
    }
}

void precond_state_free(PrecondPlan* p) { delete p; }

}  // namespace qg

using namespace qg;

extern "C" int qg_precond_apply(qg_ham* h, int kind, double shift, int dtype, int nbands, void* r, int64_t ld,
                                void* stream) {
    return guarded([&] {
        QG_REQUIRE(h && r && nbands >= 0, "bad arguments");
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        if (nbands == 0) return;
        precond_apply(h, kind, shift, dtype, nbands, r, ld, as_stream(stream));
    });
}
