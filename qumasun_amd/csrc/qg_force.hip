// qg_force.hip — nonlocal pseudopotential energy and forces of a block of occupied states on
// one GPU, and the 8th-order finite-difference gradient they take.
//
// Reference semantics (QUMASUN source/):
//   gradient  d/dx p = coef / (840 h) [672 (p+1 - p-1) - 168 (p+2 - p-2) + 32 (p+3 - p-3)
//             - 3 (p+4 - p-4)], periodic, per axis                  GridGradient8th.h:17-160
//   energy    E_nl = sum_s occ_s sum_{a,j} E_j c_saj^2, c = dV sum beta psi, states with
//             occ > 0 only                     qumasun_mpi.energy.h:349-375,
//                                              PseudoPotOperator_mpi_v2.h:573-644
//   forces    F_a += sum_j E_j c_saj d_saj (per axis), d = dV sum beta (-2 occ_s d psi_s / dx)
//                                              qumasun_mpi.force.h:226-258,
//                                              PseudoPotOperator_mpi_v2.h:905-1013
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>
#include "qg_internal.hpp"

namespace qg {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxRun = 256;                       // bands per projection call
constexpr int64_t kGradBytes = int64_t(4) << 30;   // gradient scratch of one run

// Tile of the gradient kernel: kTX x kTY points of an xy plane, one per thread, marched over z.
constexpr int kTX = 64, kTY = 4;
constexpr int kHX = kTX + 8, kHY = kTY + 8;   // the tile with its 4-point halo
constexpr int kHalo = kHX * kHY;
constexpr int kLoadsPT = (kHalo + kThreads - 1) / kThreads;
constexpr int kXcds = 8;

__device__ __forceinline__ int wrap_any(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// out_d[b][i] = coef * D_d psi[b][i] for the three axes d (periodic 8th-order central
// differences in the states' precision).  A workgroup owns a kTX x kTY column tile of one band and
// walks the planes z' = -4 .. nz+3 (periodic images at both ends).  Each plane of the tile and
// its halo is read from global memory once, two planes ahead of its use, into one of two LDS
// buffers (one barrier per plane); the x and y differences of plane z' are formed from LDS at
// once, the z difference of plane z'-4 from a register queue of this thread's nine column
// values.  The flat workgroup index is remapped so that each XCD works on a contiguous range
// of (band, tile) pairs: tiles that share halo rows run on the same XCD and meet in its L2.
template <class T>
__global__ __launch_bounds__(kThreads) void gradient8_kernel(const T* __restrict__ psi, int64_t ld, int nx, int ny,
                                                              int nz, int tiles_x, int ntiles, int nwork,
                                                              int per_xcd, double cx, double cy, double cz,
                                                              T* __restrict__ gx, T* __restrict__ gy,
                                                              T* __restrict__ gz, int64_t ld_out) {
    // the tile with its halo, padded to one slot per load so that loads and stores carry no
    // conditions (the compiler can then keep the next planes' loads in flight across them)
    __shared__ T buf[2][kLoadsPT * kThreads];
    const int lin = blockIdx.x;
    const int work = (lin % kXcds) * per_xcd + lin / kXcds;
    if (work >= nwork) return;   // padding of the remap: the whole workgroup leaves
    const int band = work / ntiles;
    const int tile = work - band * ntiles;
    const int x0 = (tile % tiles_x) * kTX, y0 = (tile / tiles_x) * kTY;
    const T* p = psi + (int64_t)band * ld;
    const int64_t sxy = (int64_t)nx * ny;
    // the global xy offsets of this thread's halo loads (fixed over z)
    int64_t src[kLoadsPT];
#pragma unroll
    for (int k = 0; k < kLoadsPT; ++k) {
        const int e = threadIdx.x + k * kThreads;
        const int hy = e / kHX, hx = e - hy * kHX;
        src[k] = e < kHalo ? (int64_t)wrap_any(x0 - 4 + hx, nx) + (int64_t)nx * wrap_any(y0 - 4 + hy, ny) : 0;
    }
    const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
    const bool active = x0 + tx < nx && y0 + ty < ny;
    const int c = (ty + 4) * kHX + tx + 4;   // this thread's point in a plane buffer
    const int64_t ob = (int64_t)band * ld_out + (x0 + tx) + (int64_t)nx * (y0 + ty);
    // planes -4 .. nz+3 are staged; the step count is padded to even (the pad step writes
    // nothing) so that the loop below needs no condition
    const int zend = nz + 4 + (nz & 1);
    using A = T;                              // arithmetic in the states' precision
    const A ax = (A)cx, ay = (A)cy, az = (A)cz;
    T q[9];                                   // column values of planes z'-8 .. z'
#pragma unroll
    for (int k = 0; k < 9; ++k) q[k] = T(0);

    auto fetch = [&](int z, T (&v)[kLoadsPT]) {
        const int64_t zo = sxy * wrap_any(z, nz);
#pragma unroll
        for (int k = 0; k < kLoadsPT; ++k) v[k] = p[zo + src[k]];
    };
    // plane z' arrives in v: stage it, fetch plane z'+2 into v, difference what is ready
    auto step = [&](int z, T (&v)[kLoadsPT]) {
        T* b = buf[z & 1];
#pragma unroll
        for (int k = 0; k < kLoadsPT; ++k) b[threadIdx.x + k * kThreads] = v[k];
        fetch(z + 2, v);   // unconditional (periodic, in bounds): the waits stay counted
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = q[k + 1];
        q[8] = b[c];
        if (!active) return;
        if (z >= 0 && z < nz) {
            A dx = A(672) * ((A)b[c + 1] - (A)b[c - 1]);
            A dy = A(672) * ((A)b[c + kHX] - (A)b[c - kHX]);
            dx += A(-168) * ((A)b[c + 2] - (A)b[c - 2]);
            dy += A(-168) * ((A)b[c + 2 * kHX] - (A)b[c - 2 * kHX]);
            dx += A(32) * ((A)b[c + 3] - (A)b[c - 3]);
            dy += A(32) * ((A)b[c + 3 * kHX] - (A)b[c - 3 * kHX]);
            dx += A(-3) * ((A)b[c + 4] - (A)b[c - 4]);
            dy += A(-3) * ((A)b[c + 4 * kHX] - (A)b[c - 4 * kHX]);
            gx[ob + sxy * z] = (T)(ax * dx);
            gy[ob + sxy * z] = (T)(ay * dy);
        }
        if (z >= 4 && z - 4 < nz) {
            A dz = A(672) * ((A)q[5] - (A)q[3]);
            dz += A(-168) * ((A)q[6] - (A)q[2]);
            dz += A(32) * ((A)q[7] - (A)q[1]);
            dz += A(-3) * ((A)q[8] - (A)q[0]);
            gz[ob + sxy * (z - 4)] = (T)(az * dz);
        }
    };
    T va[kLoadsPT], vb[kLoadsPT];
    fetch(-4, va);
    fetch(-3, vb);
    for (int z = -4; z < zend; z += 2) {
        step(z, va);
        step(z + 1, vb);
    }
}

// a status of the C ABI called from inside: failures propagate with their message
void check_rc(int rc) {
    if (rc != QG_OK) throw Error{rc, std::string(qg_last_error())};
}

void gradient8(const qg_grid& g, int dtype, int nbands, const void* psi, int64_t ld, void* gx, void* gy, void* gz,
               int64_t ld_out, double coef, hipStream_t s) {
    const double cx = coef / (840.0 * g.h[0]), cy = coef / (840.0 * g.h[1]), cz = coef / (840.0 * g.h[2]);
    const int tiles_x = (g.n[0] + kTX - 1) / kTX, tiles_y = (g.n[1] + kTY - 1) / kTY;
    const int64_t ntiles = (int64_t)tiles_x * tiles_y;
    const int64_t nwork = ntiles * nbands;
    QG_REQUIRE(nwork < (int64_t(1) << 31) - kXcds, "too many gradient tiles in one launch");
    const int64_t per_xcd = (nwork + kXcds - 1) / kXcds;
    const dim3 grid((unsigned)(per_xcd * kXcds));
    if (dtype == QG_F64)
        hipLaunchKernelGGL(gradient8_kernel<double>, grid, dim3(kThreads), 0, s, (const double*)psi, ld, g.n[0],
                           g.n[1], g.n[2], tiles_x, (int)ntiles, (int)nwork, (int)per_xcd, cx, cy, cz, (double*)gx,
                           (double*)gy, (double*)gz, ld_out);
    else
        hipLaunchKernelGGL(gradient8_kernel<float>, grid, dim3(kThreads), 0, s, (const float*)psi, ld, g.n[0], g.n[1],
                           g.n[2], tiles_x, (int)ntiles, (int)nwork, (int)per_xcd, cx, cy, cz, (float*)gx,
                           (float*)gy, (float*)gz, ld_out);
    QG_HIP(hipGetLastError());
}

}  // namespace
}  // namespace qg

using namespace qg;

extern "C" {

int qg_gradient8_apply(const qg_grid* g, int dtype, int nbands, const void* psi, int64_t ld, void* gx, void* gy,
                       void* gz, int64_t ld_out, double coef, void* stream) {
    return guarded([&] {
        QG_REQUIRE(g, "null grid");
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        QG_REQUIRE(nbands >= 0 && ld >= g->ng() && ld_out >= g->ng(), "bad band block");
        if (nbands == 0) return;
        QG_REQUIRE(psi && gx && gy && gz, "null pointer");
        hipStream_t s = as_stream(stream);
        gradient8(*g, dtype, nbands, psi, ld, gx, gy, gz, ld_out, coef, s);
        QG_HIP(hipStreamSynchronize(s));
    });
}

int qg_nonlocal_energy_force(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld,
                             const double* occ_host, double* energy, double* forces_host, void* stream) {
    return guarded([&] {
        QG_REQUIRE(nl, "null nonlocal handle");
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        QG_REQUIRE(nbands >= 0 && ld >= nl->gcopy.ng(), "bad band block");
        QG_REQUIRE(nbands == 0 || (psi && occ_host), "null pointer");
        const int natoms = nl->geo.natoms;
        if (forces_host)
            for (int i = 0; i < 3 * natoms; ++i) forces_host[i] = 0.0;
        double ev = 0.0;
        const int ncoef = nl->geo.ncoef;
        if (nbands == 0 || ncoef == 0) {
            if (energy) *energy = 0.0;
            return;
        }
        hipStream_t s = as_stream(stream);
        const qg_grid& g = nl->gcopy;
        const int64_t ng = g.ng();
        const size_t es = dtype == QG_F32 ? 4 : 8;
        // bands per run: the three gradient blocks of a run take at most kGradBytes
        const int64_t per_band = 3 * ng * (int64_t)es;
        const int chunk = (int)std::max<int64_t>(
            1, std::min<int64_t>({(int64_t)kMaxRun, (int64_t)nbands, kGradBytes / per_band}));
        // scratch kept in the handle (grown, never shrunk): coefficients of psi [chunk][ncoef],
        // then of the gradients [3][nb][ncoef]; the gradient blocks [3][nb][ng]
        nl->force_coef.alloc((size_t)4 * chunk * ncoef);
        if (forces_host) nl->force_grad.alloc((size_t)3 * chunk * ng * es);
        double* c = nl->force_coef.p;
        double* d = c + (size_t)chunk * ncoef;
        std::vector<double> hc((size_t)chunk * ncoef), hd((size_t)3 * chunk * ncoef);
        const std::vector<int>& aoff = nl->geo.atom_coef_off;
        const std::vector<double>& ecoef = nl->geo.coef_energy;
        // the occupied states (occ > 0), in runs of consecutive bands
        std::vector<int> occupied;
        for (int b = 0; b < nbands; ++b)
            if (occ_host[b] > 0.0) occupied.push_back(b);
        for (size_t k0 = 0; k0 < occupied.size();) {
            size_t k1 = k0 + 1;
            while (k1 < occupied.size() && (int)(k1 - k0) < chunk && occupied[k1] == occupied[k1 - 1] + 1) ++k1;
            const int b0 = occupied[k0], nb = (int)(k1 - k0);
            const unsigned char* p = (const unsigned char*)psi + (size_t)b0 * ld * es;
            check_rc(qg_nonlocal_project(nl, dtype, nb, p, ld, c, s));
            if (forces_host) {
                // the three gradient blocks back to back: one projection of 3 nb bands
                unsigned char* gx = nl->force_grad.p;
                unsigned char* gy = gx + (size_t)nb * ng * es;
                unsigned char* gz = gy + (size_t)nb * ng * es;
                gradient8(g, dtype, nb, p, ld, gx, gy, gz, ng, 1.0, s);
                check_rc(qg_nonlocal_project(nl, dtype, 3 * nb, gx, ng, d, s));
                QG_HIP(hipMemcpyAsync(hd.data(), d, (size_t)3 * nb * ncoef * 8, hipMemcpyDeviceToHost, s));
            }
            QG_HIP(hipMemcpyAsync(hc.data(), c, (size_t)nb * ncoef * 8, hipMemcpyDeviceToHost, s));
            QG_HIP(hipStreamSynchronize(s));
            for (int k = 0; k < nb; ++k) {
                const double occ = occ_host[b0 + k];
                const double* cb = hc.data() + (size_t)k * ncoef;
                double e = 0.0;
                for (int a = 0; a < natoms; ++a) {
                    double f[3] = {0.0, 0.0, 0.0};
                    for (int j = aoff[a]; j < aoff[a + 1]; ++j) {
                        e += cb[j] * cb[j] * ecoef[j];
                        if (forces_host)
                            for (int ax = 0; ax < 3; ++ax)
                                f[ax] += (-2.0 * occ * hd[((size_t)ax * nb + k) * ncoef + j]) * cb[j] * ecoef[j];
                    }
                    if (forces_host)
                        for (int ax = 0; ax < 3; ++ax) forces_host[3 * a + ax] += f[ax];
                }
                ev += e * occ;
            }
            k0 = k1;
        }
        if (energy) *energy = ev;
    });
}

}  // extern "C"
